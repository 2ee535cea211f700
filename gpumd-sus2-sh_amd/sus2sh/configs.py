"""The five benchmark / parity workloads, built from seeds.

C1  Cu-Zr l2k2 (K=2, 8 radial basis functions, rc 5.0), 10x10x10 fcc a=3.80 (4,000 atoms),
    50/50 species, jitter 0.05 A; one evaluation + 100 NVE steps at 1 fs from 300 K.
C2  Cu-Zr l3k3 (rb 8, rc 5.0) on 20x20x20 fcc a=3.85 (32,000 atoms), 64/36, jitter 0.08;
    1,000 NVE steps at 0.5 fs: energy drift and FP32-vs-FP64 force error.
C3  Cu-Zr l3k3 at 1,024,000 atoms (80x80x40 fcc, a=3.85, 50/50, jitter 0.08): the
    fixed-cell force loop measured by bench.py.
C4  Cu-Zr-Al l4k4 (rb 10, RBChebyshev_sss_rational, rc 5.5), 256,000 atoms at random
    positions (density 0.060 A^-3, min separation 2.0 A, 45/45/10): FP32 forward + adjoint
    for forces, virial and dE/dtheta.
C5  4,096 periodic Cu-Zr cells with U[8,256] atoms, triclinic (density 0.065 A^-3, angles
    75-105 deg), random species; l3k3; one batched evaluation with dE/dtheta.

Model choices shared by all configs: body order 5 with sh_body_l_max = {L,L,L,L},
coefficients N(0, 0.1) from the config's seed (see synthetic_model).  The MD configs (C1, C2)
add the repulsive pair core (pair_core=5 eV, scaling 0.05) so that NVE stays bounded;
C3-C5 use scaling 0.5 (C4: 0.3) so that RMS forces are O(0.3-1) eV/A.
Positions in C5 occupy random sites of a jittered sub-grid of each cell (minimum separation
about 1.7 A) instead of unconstrained uniform positions, which would put atoms on top of
each other.
"""
import numpy as np

from . import structures as st
from .model_gen import synthetic_model

MASSES = {"Cu": 63.546, "Zr": 91.224, "Al": 26.9815}


def model_c1():
    return synthetic_model(2, 2, 5, species_count=2, rb_size=8, max_dist=5.0, seed=1003,
                           scaling=0.05, pair_core=5.0)


def structure_c1(n=10):
    pos, types, cell = st.fcc_alloy(n, n, n, 3.80, [0.5, 0.5], 0.05, 1001, 1002)
    return pos, types, cell


def model_c2(md=True):
    if md:
        return synthetic_model(3, 3, 5, species_count=2, rb_size=8, max_dist=5.0, seed=2003,
                               scaling=0.05, pair_core=5.0)
    return synthetic_model(3, 3, 5, species_count=2, rb_size=8, max_dist=5.0, seed=2003,
                           scaling=0.5)


def structure_c2(n=20):
    return st.fcc_alloy(n, n, n, 3.85, [0.64, 0.36], 0.08, 2001, 2002)


def model_c3():
    return synthetic_model(3, 3, 5, body_l_max=(3, 3, 3, 3), species_count=2, rb_size=8,
                           max_dist=5.0, seed=3003, scaling=0.5)


def structure_c3(nx=80, ny=80, nz=40):
    return st.fcc_alloy(nx, ny, nz, 3.85, [0.5, 0.5], 0.08, 3001, 3002)


def model_c4():
    return synthetic_model(4, 4, 5, species_count=3, rb_size=10, max_dist=5.5, seed=4003,
                           radial_basis_type="RBChebyshev_sss_rational", scaling=0.3)


def structure_c4(n=256000):
    return random_min_separation(n, 0.060, 2.0, [0.45, 0.45, 0.10], 4001)


def model_c5():
    return synthetic_model(3, 3, 5, species_count=2, rb_size=8, max_dist=5.0, seed=5003,
                           scaling=0.5)


def random_min_separation(n, density, min_sep, fractions, seed, batch=None):
    """Random sequential addition in a periodic cube of side (n/density)^(1/3): candidates
    closer than min_sep (minimum image) to an accepted atom, or to an earlier candidate of
    the same batch, are rejected."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    side = (n / density) ** (1.0 / 3.0)
    batch = batch or max(512, n // 100)
    pts = np.zeros((0, 3))
    while len(pts) < n:
        cand = rng.uniform(0.0, side, size=(batch, 3))
        if len(pts):
            tree = cKDTree(pts, boxsize=side)
            d, _ = tree.query(cand, k=1, distance_upper_bound=min_sep)
            cand = cand[~np.isfinite(d)]
        if len(cand) > 1:
            pairs = cKDTree(cand, boxsize=side).query_pairs(min_sep, output_type="ndarray")
            drop = np.zeros(len(cand), dtype=bool)
            for a, b in sorted(map(tuple, pairs)):
                if not drop[a]:
                    drop[b] = True
            cand = cand[~drop]
        pts = np.concatenate([pts, cand[: n - len(pts)]])
    types = st.assign_species(n, fractions, seed + 1)
    return pts.T.copy(), types, np.diag([side, side, side])


def batch_c5(count=4096, nmin=8, nmax=256, seed=5001, density=0.065):
    """Triclinic cells with random atom counts, species and grid-site positions.
    Returns (cells, positions [3, N], types [N], atom_offsets [count + 1])."""
    rng = np.random.default_rng(seed)
    cells, pos, types, offs = [], [], [], [0]
    for _ in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        cell = st.triclinic_cell(n, density, rng)
        m = int(np.ceil(n ** (1.0 / 3.0))) + 1
        sites = rng.choice(m ** 3, size=n, replace=False)
        frac = np.stack([sites % m, (sites // m) % m, sites // (m * m)], axis=1).astype(float)
        frac = (frac + 0.5 + rng.uniform(-0.15, 0.15, size=frac.shape)) / m
        cells.append(cell)
        pos.append((frac @ cell).T)
        types.append(rng.integers(0, 2, size=n).astype(np.int32))
        offs.append(offs[-1] + n)
    return cells, np.ascontiguousarray(np.concatenate(pos, axis=1)), np.concatenate(types), \
        np.array(offs, dtype=np.int32)
