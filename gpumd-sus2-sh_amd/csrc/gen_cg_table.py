#!/usr/bin/env python3
"""Writes cg_table.hpp: the nonzero real Clebsch-Gordan couplings <l1 m1; l2 m2 | L M> of the
real solid harmonics for l1, l2 <= 3 (so L <= 6), as compile-time tables for the block graph
kernel and its host builder.

The coupling is the one of the standard graph generator (src/force/sus2_sh.cu:1090-1218,
restated in csrc/model.cpp: real_cg): the complex Clebsch-Gordan coefficient (Racah's
formula) between the unitary maps from complex to real harmonics (no Condon-Shortley phase),
times the phase i^(l1 + l2 - L).  Run it from this directory:  python3 gen_cg_table.py
(an optional argument names another output file).
"""
import math
import sys

LMAX = 3


def fact(n):
    return float(math.factorial(n))


def cg(j1, m1, j2, m2, j, m):
    if m != m1 + m2 or not (abs(j1 - j2) <= j <= j1 + j2):
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m) > j:
        return 0.0
    norm = math.sqrt((2 * j + 1) * fact(j1 + j2 - j) * fact(j1 - j2 + j) * fact(j2 - j1 + j) /
                     fact(j1 + j2 + j + 1))
    mfac = math.sqrt(fact(j1 + m1) * fact(j1 - m1) * fact(j2 + m2) * fact(j2 - m2) *
                     fact(j + m) * fact(j - m))
    s = 0.0
    for k in range(max(0, j2 - j - m1, j1 + m2 - j), min(j1 + j2 - j, j1 - m1, j2 + m2) + 1):
        den = (fact(k) * fact(j1 + j2 - j - k) * fact(j1 - m1 - k) * fact(j2 + m2 - k) *
               fact(j - j2 + m1 + k) * fact(j - j1 - m2 + k))
        s += (-1.0 if k & 1 else 1.0) / den
    return norm * mfac * s


def u(real_m, complex_m):
    """Row real_m of the map from complex Y_l^m to the real harmonic (no Condon-Shortley)."""
    h = 1.0 / math.sqrt(2.0)
    if real_m == 0:
        return 1.0 + 0j if complex_m == 0 else 0j
    a = abs(real_m)
    if abs(complex_m) != a:
        return 0j
    sgn = -1.0 if a & 1 else 1.0
    if real_m > 0:
        return complex(h * sgn, 0.0) if complex_m == a else complex(h, 0.0)
    return complex(0.0, -h * sgn) if complex_m == a else complex(0.0, h)


def real_cg(l1, m1, l2, m2, L, M):
    acc = 0j
    for cM in range(-L, L + 1):
        uo = u(M, cM).conjugate()
        if uo == 0:
            continue
        for c1 in range(-l1, l1 + 1):
            u1 = u(m1, c1)
            if u1 == 0:
                continue
            for c2 in range(-l2, l2 + 1):
                u2 = u(m2, c2)
                if u2 == 0:
                    continue
                c = cg(l1, c1, l2, c2, L, cM)
                if abs(c) < 1e-14:
                    continue
                acc += uo * c * u1 * u2
    acc *= [1, 1j, -1, -1j][(l1 + l2 - L) & 3]
    assert abs(acc.imag) < 1e-10
    return acc.real


def main(path="cg_table.hpp"):
    types = [(l1, l2, L) for l1 in range(LMAX + 1) for l2 in range(LMAX + 1)
             for L in range(abs(l1 - l2), l1 + l2 + 1)]
    out = []
    w = out.append
    w("// Generated by gen_cg_table.py -- do not edit.  Nonzero real Clebsch-Gordan couplings")
    w("// <l1 m1; l2 m2 | L M> of the real solid harmonics (the standard graph generator's")
    w("// coupling, src/force/sus2_sh.cu:1090-1218) for l1, l2 <= %d: block type t couples a" % LMAX)
    w("// tensor of degree kCgL1[t] with one of degree kCgL2[t] into degree kCgLL[t]; entry")
    w("// {M + L, m1 + l1, m2 + l2, c} of CgTable<t> adds c a[m1] b[m2] to component M.")
    w("#pragma once")
    w("")
    w("namespace sus2sh {")
    w("")
    w("constexpr int kCgLMax = %d;" % LMAX)
    w("constexpr int kCgTypes = %d;" % len(types))
    w("constexpr int kCgL1[kCgTypes] = {%s};" % ", ".join(str(t[0]) for t in types))
    w("constexpr int kCgL2[kCgTypes] = {%s};" % ", ".join(str(t[1]) for t in types))
    w("constexpr int kCgLL[kCgTypes] = {%s};" % ", ".join(str(t[2]) for t in types))
    w("")
    w("struct CgEntry {")
    w("  int M, m1, m2;")
    w("  double c;")
    w("};")
    w("")
    w("template <int T>")
    w("struct CgTable;")
    counts = []
    for t, (l1, l2, L) in enumerate(types):
        ent = []
        for M in range(-L, L + 1):
            for m1 in range(-l1, l1 + 1):
                for m2 in range(-l2, l2 + 1):
                    c = real_cg(l1, m1, l2, m2, L, M)
                    if abs(c) > 1e-13:
                        ent.append((M + L, m1 + l1, m2 + l2, c))
        counts.append(len(ent))
        w("template <>")
        w("struct CgTable<%d> {  // (%d, %d) -> %d" % (t, l1, l2, L))
        w("  static constexpr int n = %d;" % len(ent))
        w("  static constexpr CgEntry e[n] = {")
        for e in ent:
            w("    {%d, %d, %d, %s}," % (e[0], e[1], e[2], repr(float(e[3]))))
        w("  };")
        w("};")
    w("")
    w("constexpr int kCgEntries[kCgTypes] = {%s};" % ", ".join(str(c) for c in counts))
    w("")
    w("// X(t, l1, l2, L) for every block type")
    w("#define SUS2SH_CG_TYPES(X) \\")
    w(" \\\n".join("  X(%d, %d, %d, %d)" % (t, *types[t]) for t in range(len(types))))
    w("")
    w("}  // namespace sus2sh")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main(*sys.argv[1:2])
