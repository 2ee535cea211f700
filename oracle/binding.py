"""ctypes binding of the FP64 CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py use this module,
always as the checker, never as the engine.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_sus2sh.so")
_lib = None


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def build_native(out_path):
    """Compile the oracle with -O3 -march=native on this machine into out_path (for the CPU
    baseline, which is timed on the host that runs it)."""
    subprocess.run(["make", "-C", _HERE, "-s", "native", f"NATIVE_OUT={out_path}"], check=True)


def use_library(path):
    """Load the oracle from another build of the same source (e.g. build_native's output);
    every later call in this process uses it."""
    global _lib
    _lib = _configure(ctypes.CDLL(path))
    return _lib


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    _lib = _configure(ctypes.CDLL(LIB_PATH))
    return _lib


def _configure(L):
    P = ctypes.POINTER
    c_int, c_double, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    L.oracle_load.argtypes = [ctypes.c_char_p]
    L.oracle_load.restype = c_void_p
    L.oracle_free.argtypes = [c_void_p]
    L.oracle_free.restype = None
    L.oracle_dims.argtypes = [c_void_p, P(c_int)]
    L.oracle_dims.restype = None
    L.oracle_cutoff.argtypes = [c_void_p]
    L.oracle_cutoff.restype = c_double
    L.oracle_param_count.argtypes = [c_void_p]
    L.oracle_param_count.restype = c_int
    L.oracle_compute.argtypes = [c_void_p, c_int, P(c_int), P(c_double), P(c_double), P(c_int),
                                 c_int, P(c_double), P(c_double), P(c_double), P(c_double)]
    L.oracle_compute.restype = c_int
    L.oracle_graph.argtypes = [c_void_p, c_int, P(c_double), P(c_double), P(c_double)]
    L.oracle_graph.restype = c_int
    L.oracle_ylm.argtypes = [c_int, c_double, c_double, c_double, P(c_double), P(c_double)]
    L.oracle_ylm.restype = None
    L.oracle_radial.argtypes = [c_void_p, c_int, c_int, c_double, P(c_double), P(c_double)]
    L.oracle_radial.restype = None
    L.oracle_neighbor_counts.argtypes = [c_int, P(c_double), P(c_double), P(c_int), c_double,
                                         c_int, P(c_int)]
    L.oracle_neighbor_counts.restype = c_int
    L.oracle_num_threads.argtypes = []
    L.oracle_num_threads.restype = c_int
    L.oracle_set_num_threads.argtypes = [c_int]
    L.oracle_set_num_threads.restype = None
    L.oracle_zbl_pair.argtypes = [c_int, c_int, c_double, c_double, c_double, P(c_double),
                                  P(c_double)]
    L.oracle_zbl_pair.restype = None
    L.oracle_set_zbl.argtypes = [c_void_p, P(c_int), P(c_double), P(c_double)]
    L.oracle_set_zbl.restype = c_int
    L.oracle_zbl_enabled.argtypes = [c_void_p]
    L.oracle_zbl_enabled.restype = c_int
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def ylm(lmax, vec):
    """Real Y_lm values [(lmax+1)^2] and gradients [(lmax+1)^2, 3] at a displacement."""
    n = (lmax + 1) ** 2
    Y = np.zeros(n)
    dY = np.zeros(3 * n)
    lib().oracle_ylm(lmax, float(vec[0]), float(vec[1]), float(vec[2]), _dp(Y), _dp(dY))
    return Y, dY.reshape(n, 3)


def neighbor_counts(cell, positions, rc, pbc=(1, 1, 1), brute=True):
    from_box = _box(cell)
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
    n = pos.size // 3
    out = np.zeros(n, dtype=np.int32)
    pb = np.asarray(pbc, dtype=np.int32)
    lib().oracle_neighbor_counts(n, _dp(pos), _dp(from_box), _ip(pb), float(rc),
                                 0 if brute else 1, _ip(out))
    return out


def zbl_pair(zi, zj, r, inner, outer):
    """(phi, dphi/dr) of one ZBL pair (eV, eV/A)."""
    e, de = ctypes.c_double(), ctypes.c_double()
    lib().oracle_zbl_pair(int(zi), int(zj), float(r), float(inner), float(outer),
                          ctypes.byref(e), ctypes.byref(de))
    return e.value, de.value


def _box(cell):
    H = np.asarray(cell, dtype=np.float64).T
    return np.ascontiguousarray(np.concatenate([H.reshape(-1), np.linalg.inv(H).reshape(-1)]))


class Oracle:
    """FP64 restatement of the SUS2-SH path for one `.mtp` model."""

    def __init__(self, mtp_path):
        self._m = lib().oracle_load(str(mtp_path).encode())
        if not self._m:
            raise ValueError(f"oracle could not parse {mtp_path}")
        d = (ctypes.c_int * 12)()
        lib().oracle_dims(self._m, d)
        keys = ("S", "L", "K", "body", "rb", "Mr", "A", "B", "Q", "T", "kind", "gate_scalars")
        self.dims = dict(zip(keys, list(d)))
        self.rc = lib().oracle_cutoff(self._m)
        self.param_count = lib().oracle_param_count(self._m)

    def __del__(self):
        if getattr(self, "_m", None):
            lib().oracle_free(self._m)
            self._m = None

    @property
    def zbl_enabled(self):
        return bool(lib().oracle_zbl_enabled(self._m))

    def set_zbl(self, atomic_numbers, inner, outer):
        """ZBL model: atomic number per species and [S, S] pair cutoffs (A)."""
        S = self.dims["S"]
        z = np.ascontiguousarray(atomic_numbers, dtype=np.int32)
        a = np.ascontiguousarray(np.broadcast_to(inner, (S, S)), dtype=np.float64)
        b = np.ascontiguousarray(np.broadcast_to(outer, (S, S)), dtype=np.float64)
        if z.size != S or lib().oracle_set_zbl(self._m, _ip(z), _dp(a), _dp(b)) != 0:
            raise ValueError("set_zbl: one atomic number per species (at most 16 species)")

    def radial(self, ti, tj, r):
        Mr = self.dims["Mr"]
        R = np.zeros(Mr)
        dR = np.zeros(Mr)
        lib().oracle_radial(self._m, ti, tj, float(r), _dp(R), _dp(dR))
        return R, dR

    def graph(self, basics):
        """Graph part alone: basics [n, B] -> (E' [n], dE'/dbasic [n, B])."""
        b = np.ascontiguousarray(np.atleast_2d(basics), dtype=np.float64)
        n = b.shape[0]
        e = np.zeros(n)
        g = np.zeros_like(b)
        if lib().oracle_graph(self._m, n, _dp(b), _dp(e), _dp(g)) != 0:
            raise MemoryError("oracle_graph")
        return e, g

    def compute(self, cell, types, positions, pbc=(1, 1, 1), brute=False, param_grad=False):
        """Returns (energy[N], force[3,N], virial[9,N]) and optionally dE/dtheta."""
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        n = pos.size // 3
        t = np.ascontiguousarray(types, dtype=np.int32)
        e = np.zeros(n)
        f = np.zeros(3 * n)
        v = np.zeros(9 * n)
        pg = np.zeros(self.param_count) if param_grad else None
        pb = np.asarray(pbc, dtype=np.int32)
        box = _box(cell)
        rc = lib().oracle_compute(self._m, n, _ip(t), _dp(pos), _dp(box), _ip(pb),
                                  1 if brute else 0, _dp(e), _dp(f), _dp(v),
                                  _dp(pg) if pg is not None else None)
        if rc == -2:
            raise ValueError("oracle_compute: gate models have no dE/dtheta or ZBL here")
        if rc != 0:
            raise ValueError("oracle_compute: ZBL model without set_zbl()")
        out = (e, f.reshape(3, n), v.reshape(9, n))
        return out + (pg,) if param_grad else out
