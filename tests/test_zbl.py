"""ZBL pair term -- CPU only: the oracle's pair function, the engine's host tables, and the
oracle's energy / forces / virial with ZBL.

The oracle pair function is pinned by what the physics fixes: the Coulomb prefactor
e^2 / (4 pi eps0) from CODATA constants, the screening function's normalisation
(sum of amplitudes = 1, so phi -> Coulomb as r -> 0), the published Ziegler-Biersack-Littmark
decay constants (at the reference's screening constant 2.134563), central differences of
phi, and the switch (phi unchanged below inner, zero value and slope at outer).
The host tables replay the reference's own known-answer checks
(codex_tests/gpumd_sh_zbl_common_check.cpp: defaults, symbol mapping, H/C covalent radii,
the H-C type-wise outer cutoff) plus its error messages.
"""
import math

import numpy as np
import pytest

import sus2sh
from oracle.binding import Oracle, zbl_pair

# CODATA 2018: elementary charge (C), vacuum permittivity (F/m)
E_CHARGE = 1.602176634e-19
EPS0 = 8.8541878128e-12
COULOMB_EV_A = E_CHARGE / (4 * math.pi * EPS0) * 1e10  # e^2/(4 pi eps0) in eV A


def test_coulomb_prefactor_matches_codata():
    assert abs(COULOMB_EV_A - 14.3996454784255) < 1e-9


@pytest.mark.parametrize("zi,zj", [(1, 6), (29, 40), (40, 40), (13, 29)])
def test_short_distance_limit_is_bare_coulomb(zi, zj):
    r = 1e-7
    e, _ = zbl_pair(zi, zj, r, 0.7, 1.4)
    assert abs(e * r / (COULOMB_EV_A * zi * zj) - 1.0) < 1e-5


def test_screening_against_published_form():
    # phi(x) = 0.18175 e^-3.1998x + 0.50986 e^-0.94229x + 0.28022 e^-0.4029x + 0.02817 e^-0.20162x
    # with x = r / a, 1 / a = 2.134563 (Zi^0.23 + Zj^0.23); inside inner the switch is 1
    zi, zj, r = 29, 40, 0.35
    x = 2.134563 * (zi ** 0.23 + zj ** 0.23) * r
    phi = (0.18175 * math.exp(-3.1998 * x) + 0.50986 * math.exp(-0.94229 * x) +
           0.28022 * math.exp(-0.4029 * x) + 0.02817 * math.exp(-0.20162 * x))
    e, _ = zbl_pair(zi, zj, r, 0.7, 1.4)
    assert abs(e - COULOMB_EV_A * zi * zj * phi / r) < 1e-10 * abs(e)


@pytest.mark.parametrize("r", [0.3, 0.69, 0.75, 1.0, 1.3, 1.39])
def test_derivative_central_differences(r):
    h = 1e-6
    _, de = zbl_pair(29, 40, r, 0.7, 1.4)
    ep, _ = zbl_pair(29, 40, r + h, 0.7, 1.4)
    em, _ = zbl_pair(29, 40, r - h, 0.7, 1.4)
    assert abs(de - (ep - em) / (2 * h)) < 1e-6 * max(1.0, abs(de))


def test_switch_endpoints():
    zi, zj = 29, 40
    # below inner the switch is 1: independent of the (inner, outer) pair
    a, _ = zbl_pair(zi, zj, 0.5, 0.6, 1.4)
    b, _ = zbl_pair(zi, zj, 0.5, 0.9, 2.0)
    assert a == b
    # value and slope vanish at outer; nothing beyond
    e1, d1 = zbl_pair(zi, zj, 1.4 - 1e-7, 0.7, 1.4)
    assert abs(e1) < 1e-10 and abs(d1) < 1e-4
    assert zbl_pair(zi, zj, 1.4, 0.7, 1.4) == (0.0, 0.0)
    assert zbl_pair(zi, zj, 2.0, 0.7, 1.4) == (0.0, 0.0)
    # at mid-switch the factor is 1/2
    inner, outer, r = 0.7, 1.4, 1.05
    e, _ = zbl_pair(zi, zj, r, inner, outer)
    e_free, _ = zbl_pair(zi, zj, r, 2.0, 3.0)  # unswitched
    assert abs(e - 0.5 * e_free) < 1e-12 * abs(e_free)


def _zbl_model(model_dir, name, extra, species_count=2):
    import os
    spec = sus2sh.synthetic_model(2, 2, 5, species_count=species_count, rb_size=8,
                                  max_dist=5.0, seed=31, scaling=0.05)
    spec.extra = dict(extra)
    return sus2sh.write_mtp(spec, os.path.join(str(model_dir), name + ".mtp"))


def test_host_tables_replay_reference_known_answers(model_dir):
    # defaults 0.7 / 1.4 when only zbl_enabled is given
    path = _zbl_model(model_dir, "zbl_defaults", {"zbl_enabled": True})
    z, inner, outer = sus2sh.zbl_tables(path, ["Cu", "Zr"])
    assert list(z) == [29, 40]
    assert np.all(inner == 0.7) and np.all(outer == 1.4)
    # symbol mapping H, C, Cl, Co
    for sym, num in (("H", 1), ("C", 6), ("Cl", 17), ("Co", 27)):
        z, _, _ = sus2sh.zbl_tables(path, [sym, "Zr"])
        assert z[0] == num
    # type-wise: H-C outer = 0.7 (0.426667 + 1.0), inner 0; capped by the global outer
    path = _zbl_model(model_dir, "zbl_typewise",
                      {"zbl_enabled": True, "zbl_typewise_cutoff_factor": 0.7})
    z, inner, outer = sus2sh.zbl_tables(path, ["H", "C"])
    assert abs(outer[0, 1] - 0.7 * (0.426667 + 1.0)) < 1e-14
    assert outer[0, 1] == outer[1, 0]
    assert np.all(inner == 0.0)
    assert abs(outer[1, 1] - min(1.4, 0.7 * 2.0)) < 1e-14
    # file metadata overrides the symbols
    path = _zbl_model(model_dir, "zbl_meta",
                      {"zbl_enabled": True, "zbl_atomic_numbers": [13, 26]})
    z, _, _ = sus2sh.zbl_tables(path, ["Cu", "Zr"])
    assert list(z) == [13, 26]


@pytest.mark.parametrize("case,extra,symbols,message", [
    (0, {"zbl_enabled": True}, ["Cu", "Xx"], "cannot map element symbol"),
    (1, {"zbl_enabled": True}, ["Cu"], "requires zbl_atomic_numbers metadata"),
    (2, {"zbl_enabled": True, "zbl_inner": 1.5, "zbl_outer": 1.4}, ["Cu", "Zr"],
     "inner < outer"),
    (3, {"zbl_enabled": True, "zbl_typewise_cutoff_factor": 0.4}, ["Cu", "Zr"],
     "at least 0.5"),
    (4, {"zbl_enabled": True, "zbl_atomic_numbers": [29]}, ["Cu", "Zr"],
     "one atomic number per species"),
    (5, {"zbl_enabled": True, "zbl_atomic_numbers": [29, 0]}, ["Cu", "Zr"],
     "should be in [1, 94]"),
])
def test_host_table_errors(model_dir, case, extra, symbols, message):
    path = _zbl_model(model_dir, "zbl_err_%d" % case, extra)
    with pytest.raises(sus2sh.Sus2shError) as err:
        sus2sh.zbl_tables(path, symbols)
    assert message in err.value.message


def _cluster(n, box, dmin, seed):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.uniform(0, box, size=3)
        if all(np.linalg.norm(p - q) >= dmin for q in pts):
            pts.append(p)
    return np.array(pts).T.copy()


def test_oracle_with_zbl_forces_and_virial(model_dir):
    """Energy with the ZBL term: forces are minus its central differences, sum F = 0, and
    the ZBL part is exactly the pair sum over close pairs."""
    path = _zbl_model(model_dir, "zbl_oracle", {"zbl_enabled": True, "zbl_outer": 1.6,
                                                 "zbl_inner": 0.8})
    cell = np.diag([30.0, 30.0, 30.0])  # isolated cluster (no image within rc)
    pos = _cluster(14, 4.0, 1.0, 3)
    types = np.array([0, 1] * 7, dtype=np.int32)
    plain = Oracle(path)
    plain.set_zbl([29, 40], 0.0, 1e-9)  # pair cutoff below every distance: SH part only
    orc = Oracle(path)
    with pytest.raises(ValueError):
        orc.compute(cell, types, pos)   # ZBL model without tables
    orc.set_zbl([29, 40], 0.8, 1.6)
    e, f, v = orc.compute(cell, types, pos, brute=True)
    e0, _, _ = plain.compute(cell, types, pos, brute=True)
    pair = 0.0
    zz = np.array([29, 40])[types]
    for i in range(len(types)):
        for j in range(i + 1, len(types)):
            r = np.linalg.norm(pos[:, j] - pos[:, i])
            pair += zbl_pair(zz[i], zz[j], r, 0.8, 1.6)[0]
    assert pair > 1.0  # the cluster has pairs inside the ZBL range
    assert abs((e.sum() - e0.sum()) - pair) < 1e-10 * max(1.0, abs(pair))
    assert np.abs(f.sum(axis=1)).max() < 1e-9
    h = 1e-5
    for a, c in ((0, 0), (3, 1), (7, 2), (11, 0)):
        p1, p2 = pos.copy(), pos.copy()
        p1[c, a] += h
        p2[c, a] -= h
        fd = -(orc.compute(cell, types, p1, brute=True)[0].sum() -
               orc.compute(cell, types, p2, brute=True)[0].sum()) / (2 * h)
        assert abs(f[c, a] - fd) < 1e-6 * max(1.0, abs(fd))
    # virial: total W = sum_i r_i (x) F_i for an isolated cluster (W = -sum_edges d (x) g;
    # positive diagonal for the repulsive ZBL core)
    W = v.sum(axis=1)
    ref = np.einsum("ai,bi->ab", pos, f)
    order = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2), (1, 0), (2, 0), (2, 1)]
    for k, (a, b) in enumerate(order):
        assert abs(W[k] - ref[a, b]) < 1e-8 * max(1.0, abs(ref[a, b]))
