"""Pins of the FP64 CPU oracle (oracle/sus2sh_oracle.c) -- CPU only.

Each oracle function is checked against something that does not restate its own formula:
  Y_lm       the reference's explicit l<=4 table (src/force/sus2_sh.cu:3058-3168), the
             addition theorem, quadrature orthonormality, central differences;
  radial     T_n(cos t) = cos(n t), envelope zeros at rc, central differences;
  E/F/W      central differences of E, sum F = 0, translation / rotation / permutation
             invariance, virial = strain derivative of E, brute-force vs cell-list
             neighbours, small periodic cell vs replicated supercell;
  graph      the flat product execution is tied to the generator by rotational invariance
             (any wrong CG coefficient breaks it).
"""
import math

import numpy as np
import pytest

import sus2sh
from sus2sh import structures as st
from oracle.binding import Oracle, neighbor_counts, ylm

PI = math.pi

# (l, m, constant, polynomial) of the reference's explicit real harmonics (Y = c P / r^l)
REF_YLM = [
    (0, 0, 0.5 / math.sqrt(PI), lambda x, y, z: 1.0),
    (1, -1, 0.5 * math.sqrt(3 / PI), lambda x, y, z: y),
    (1, 0, 0.5 * math.sqrt(3 / PI), lambda x, y, z: z),
    (1, 1, 0.5 * math.sqrt(3 / PI), lambda x, y, z: x),
    (2, -2, 0.5 * math.sqrt(15 / PI), lambda x, y, z: x * y),
    (2, -1, 0.5 * math.sqrt(15 / PI), lambda x, y, z: y * z),
    (2, 0, 0.25 * math.sqrt(5 / PI), lambda x, y, z: 2 * z * z - x * x - y * y),
    (2, 1, 0.5 * math.sqrt(15 / PI), lambda x, y, z: x * z),
    (2, 2, 0.25 * math.sqrt(15 / PI), lambda x, y, z: x * x - y * y),
    (3, -3, 0.125 * math.sqrt(70 / PI), lambda x, y, z: y * (3 * x * x - y * y)),
    (3, -2, 0.5 * math.sqrt(105 / PI), lambda x, y, z: x * y * z),
    (3, -1, 0.125 * math.sqrt(42 / PI), lambda x, y, z: y * (4 * z * z - x * x - y * y)),
    (3, 0, 0.25 * math.sqrt(7 / PI), lambda x, y, z: z * (2 * z * z - 3 * x * x - 3 * y * y)),
    (3, 1, 0.125 * math.sqrt(42 / PI), lambda x, y, z: x * (4 * z * z - x * x - y * y)),
    (3, 2, 0.25 * math.sqrt(105 / PI), lambda x, y, z: z * (x * x - y * y)),
    (3, 3, 0.125 * math.sqrt(70 / PI), lambda x, y, z: x * (x * x - 3 * y * y)),
    (4, -4, 0.75 * math.sqrt(35 / PI), lambda x, y, z: x * y * (x * x - y * y)),
    (4, -3, 0.375 * math.sqrt(70 / PI), lambda x, y, z: z * y * (3 * x * x - y * y)),
    (4, -2, 0.75 * math.sqrt(5 / PI), lambda x, y, z: x * y * (6 * z * z - x * x - y * y)),
    (4, -1, 0.375 * math.sqrt(10 / PI), lambda x, y, z: y * z * (4 * z * z - 3 * (x * x + y * y))),
    (4, 0, 0.1875 / math.sqrt(PI), lambda x, y, z: 8 * z ** 4 - 24 * z * z * (x * x + y * y)
     + 3 * (x * x + y * y) ** 2),
    (4, 1, 0.375 * math.sqrt(10 / PI), lambda x, y, z: x * z * (4 * z * z - 3 * (x * x + y * y))),
    (4, 2, 0.375 * math.sqrt(5 / PI), lambda x, y, z: (x * x - y * y) * (6 * z * z - x * x - y * y)),
    (4, 3, 0.375 * math.sqrt(70 / PI), lambda x, y, z: z * x * (x * x - 3 * y * y)),
    (4, 4, 0.1875 * math.sqrt(35 / PI), lambda x, y, z: x ** 4 - 6 * x * x * y * y + y ** 4),
]


def _rand_vecs(n, seed=0):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, 3)) * rng.uniform(0.5, 4.0, size=(n, 1))


def test_ylm_matches_reference_table():
    for v in _rand_vecs(20):
        Y, _ = ylm(4, v)
        r = np.linalg.norm(v)
        for l, m, c, P in REF_YLM:
            assert Y[l * l + m + l] == pytest.approx(c * P(*v) / r ** l, rel=1e-12, abs=1e-13)


def test_ylm_addition_theorem():
    for u, v in zip(_rand_vecs(10, 1), _rand_vecs(10, 2)):
        Yu, _ = ylm(6, u)
        Yv, _ = ylm(6, v)
        c = np.dot(u, v) / (np.linalg.norm(u) * np.linalg.norm(v))
        for l in range(7):
            s = sum(Yu[l * l + m + l] * Yv[l * l + m + l] for m in range(-l, l + 1))
            Pl = np.polynomial.legendre.legval(c, [0] * l + [1])
            assert s == pytest.approx((2 * l + 1) / (4 * PI) * Pl, abs=1e-12)


def test_ylm_orthonormal_by_quadrature():
    x, w = np.polynomial.legendre.leggauss(16)
    nphi = 32
    G = np.zeros((49, 49))
    for xi, wi in zip(x, w):
        s = math.sqrt(1 - xi * xi)
        for k in range(nphi):
            ph = 2 * PI * k / nphi
            Y, _ = ylm(6, (s * math.cos(ph), s * math.sin(ph), xi))
            G += wi * (2 * PI / nphi) * np.outer(Y, Y)
    assert np.abs(G - np.eye(49)).max() < 1e-12


def test_ylm_gradient_central_differences():
    h = 1e-6
    for v in _rand_vecs(8, 3):
        _, dY = ylm(6, v)
        for c in range(3):
            vp, vm = v.copy(), v.copy()
            vp[c] += h
            vm[c] -= h
            fd = (ylm(6, vp)[0] - ylm(6, vm)[0]) / (2 * h)
            assert np.abs(fd - dY[:, c]).max() < 1e-7


@pytest.fixture(scope="module")
def l2k2(make_model):
    return make_model("oracle_l2k2", l_max=2, k_max=2, body_order=5, rb_size=8, max_dist=5.0,
                      seed=101, scaling=0.5)


@pytest.fixture(scope="module")
def l3k2_random(make_model):
    return make_model("oracle_l3k2r", l_max=3, k_max=2, body_order=4, rb_size=6, max_dist=4.5,
                      seed=102, scal_mode="random", species_count=3)


@pytest.fixture(scope="module")
def rational(make_model):
    return make_model("oracle_rat", l_max=2, k_max=2, body_order=5, rb_size=7, max_dist=5.5,
                      seed=103, radial_basis_type="RBChebyshev_sss_rational", scal_mode="random")


@pytest.fixture(scope="module")
def laguerre(make_model):
    return make_model("oracle_lag", l_max=2, k_max=2, body_order=5, rb_size=7, max_dist=5.0,
                      seed=104, radial_basis_type="RBLaguerre_log1p", scal_mode="random")


def test_laguerre_radial_closed_form(laguerre):
    """R_mu(r) = (r-rc)^2 exp(-u/2) sum_n c_n L_n(u), u = a log(1 + r/rho), with L_n the
    Laguerre polynomials of scipy.special.eval_laguerre."""
    import re
    from scipy.special import eval_laguerre
    o = Oracle(laguerre)
    text = open(laguerre).read()
    spec_scal = np.array([float(v) for v in re.search(r"scal_coeffs = \{([^}]*)\}", text).group(1).split(",")])
    rows = re.findall(r"\t\t\{([^}]*)\}", text)
    S, Mr, rb, rc = 2, 6, 7, 5.0
    scaling = float(re.search(r"scaling = (\S+)", text).group(1))
    rad = np.array([[float(v) for v in r.split(",")] for r in rows[:Mr]])
    w = rad[0, rb:]
    for ti in range(S):
        for tj in range(S):
            for r in (0.9, 2.1, 3.4, 4.7):
                R, _ = o.radial(ti, tj, r)
                for mu in range(Mr):
                    a = spec_scal[2 * mu * S * S + S * ti + tj]
                    rho = spec_scal[2 * mu * S * S + S * S + S * ti + tj]
                    u = a * math.log1p(r / rho)
                    val = sum(rad[mu, n] * eval_laguerre(n, u) for n in range(rb))
                    ref = (r - rc) ** 2 * math.exp(-0.5 * u) * scaling * w[ti] * w[tj] * val
                    assert R[mu] == pytest.approx(ref, rel=1e-11, abs=1e-13)


def test_radial_closed_form(l3k2_random):
    """R_mu(r) = (r-rc)^2 sum_n c_n cos(n acos xi), xi = tanh(a (r - s)/2)."""
    import re
    o = Oracle(l3k2_random)
    text = open(l3k2_random).read()
    spec_scal = np.array([float(v) for v in re.search(r"scal_coeffs = \{([^}]*)\}", text).group(1).split(",")])
    rows = re.findall(r"\t\t\{([^}]*)\}", text)
    S, Mr, rb, rc = 3, 8, 6, 4.5
    scaling = float(re.search(r"scaling = (\S+)", text).group(1))
    rad = np.array([[float(v) for v in r.split(",")] for r in rows[:Mr]])
    w = rad[0, rb:]
    for ti in range(S):
        for tj in range(S):
            for r in (1.8, 2.6, 3.3, 4.4):
                R, _ = o.radial(ti, tj, r)
                for mu in range(Mr):
                    a = spec_scal[2 * mu * S * S + S * ti + tj]
                    s = spec_scal[2 * mu * S * S + S * S + S * ti + tj]
                    xi = math.tanh(0.5 * a * (r - s))
                    th = math.acos(xi)
                    val = sum(rad[mu, n] * math.cos(n * th) for n in range(rb))
                    ref = (r - rc) ** 2 * scaling * w[ti] * w[tj] * val
                    assert R[mu] == pytest.approx(ref, rel=1e-11, abs=1e-13)


@pytest.mark.parametrize("fixture", ["l3k2_random", "rational", "laguerre"])
def test_radial_derivative_and_cutoff(request, fixture):
    o = Oracle(request.getfixturevalue(fixture))
    rc = o.rc
    h = 1e-6
    for ti in range(o.dims["S"]):
        for tj in range(o.dims["S"]):
            for r in (1.9, 2.7, 3.9):
                _, dR = o.radial(ti, tj, r)
                fd = (o.radial(ti, tj, r + h)[0] - o.radial(ti, tj, r - h)[0]) / (2 * h)
                assert np.abs(fd - dR).max() < 1e-7
            R, dR = o.radial(ti, tj, rc)
            assert np.abs(R).max() == 0.0 and np.abs(dR).max() == 0.0


def _struct(n=3, a=3.8, sigma=0.06, S=2, seed=1):
    return st.fcc_alloy(n, n, n, a, [1.0 / S] * S, sigma, seed, seed + 1)


@pytest.mark.parametrize("fixture", ["l2k2", "l3k2_random", "rational", "laguerre"])
def test_forces_are_energy_gradient(request, fixture):
    o = Oracle(request.getfixturevalue(fixture))
    pos, types, cell = _struct(S=o.dims["S"])
    e, f, v = o.compute(cell, types, pos)
    assert np.abs(f.sum(axis=1)).max() < 1e-10
    h = 1e-5
    rng = np.random.default_rng(0)
    for i in rng.choice(len(types), 4, replace=False):
        for c in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[c, i] += h
            pm[c, i] -= h
            fd = -(o.compute(cell, types, pp)[0].sum() - o.compute(cell, types, pm)[0].sum()) / (2 * h)
            assert fd == pytest.approx(f[c, i], abs=2e-7 * max(1.0, abs(f[c, i])))


def test_virial_is_strain_derivative(l2k2):
    """sum_i W_i^{ab} = -dE/d eps_ab under x -> (1 + eps) x (cell and atoms)."""
    o = Oracle(l2k2)
    pos, types, cell = _struct()
    _, _, v = o.compute(cell, types, pos)
    W = v.sum(axis=1)
    order = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2), (1, 0), (2, 0), (2, 1)]
    h = 1e-6
    for k, (a, b) in enumerate(order):
        def energy(eps):
            F = np.eye(3)
            F[a, b] += eps
            return o.compute(cell @ F.T, types, F @ pos)[0].sum()
        dE = (energy(h) - energy(-h)) / (2 * h)
        assert W[k] == pytest.approx(-dE, abs=1e-6 * max(1.0, abs(dE)))


def test_invariances(l3k2_random):
    o = Oracle(l3k2_random)
    pos, types, cell = _struct(S=3, seed=4)
    e, f, v = o.compute(cell, types, pos)
    # translation
    e2, f2, _ = o.compute(cell, types, pos + np.array([[0.37], [-1.2], [2.9]]))
    assert np.abs(e2 - e).max() < 1e-10 and np.abs(f2 - f).max() < 1e-10
    # rotation of cell and atoms: energies invariant, forces rotate
    rng = np.random.default_rng(9)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    e3, f3, _ = o.compute(cell @ q.T, types, q @ pos)
    assert np.abs(e3 - e).max() < 1e-10
    assert np.abs(f3 - q @ f).max() < 1e-9
    # permutation of atoms
    perm = rng.permutation(len(types))
    e4, f4, _ = o.compute(cell, types[perm], pos[:, perm])
    assert np.abs(e4 - e[perm]).max() < 1e-10 and np.abs(f4 - f[:, perm]).max() < 1e-10


def test_rotation_invariance_pins_graph(make_model):
    """Body-order-5 and -6 scalars are rotation invariant only with correct real CG (and, at
    l_max = 5, 6, Y_lm and CG of those degrees in the same basis)."""
    for name, kw in [("rot5", dict(l_max=3, k_max=2, body_order=5, rb_size=5, max_dist=4.5, seed=31)),
                     ("rot6", dict(l_max=2, k_max=2, body_order=6, rb_size=5, max_dist=4.5, seed=32)),
                     ("rot_l5", dict(l_max=5, k_max=2, body_order=5, body_l_max=(5, 5, 1, 1),
                                     rb_size=5, max_dist=4.5, seed=33)),
                     ("rot_l6", dict(l_max=6, k_max=1, body_order=5, body_l_max=(6, 6, 1, 1),
                                     rb_size=5, max_dist=4.5, seed=34))]:
        o = Oracle(make_model(name, **kw))
        rng = np.random.default_rng(3)
        pos = rng.uniform(-2.0, 2.0, size=(3, 9))
        cell = np.eye(3) * 30.0
        types = rng.integers(0, 2, size=9).astype(np.int32)
        e = o.compute(cell, types, pos, pbc=(0, 0, 0), brute=True)[0]
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        e2 = o.compute(cell, types, q @ pos, pbc=(0, 0, 0), brute=True)[0]
        assert np.abs(e - e2).max() < 1e-10 * max(1.0, np.abs(e).max())


def test_cell_list_matches_brute_force(l2k2):
    o = Oracle(l2k2)
    rng = np.random.default_rng(2)
    for n, pbc in [(3, (1, 1, 1)), (4, (1, 0, 1))]:
        pos, types, cell = _struct(n=n, seed=n)
        pos = pos + rng.normal(0, 0.2, size=pos.shape)
        a = o.compute(cell, types, pos, pbc=pbc, brute=True)
        b = o.compute(cell, types, pos, pbc=pbc, brute=False)
        for x, y in zip(a, b):
            assert np.abs(x - y).max() < 1e-11
        c1 = neighbor_counts(cell, pos, o.rc, pbc, brute=True)
        c2 = neighbor_counts(cell, pos, o.rc, pbc, brute=False)
        assert (c1 == c2).all()


def test_fcc_neighbour_shells():
    """fcc a=3.85, rc=5.0: 12 + 6 + 24 = 42 neighbours per atom."""
    pos, cell = st.fcc(4, 4, 4, 3.85)
    c = neighbor_counts(cell, pos, 5.0, brute=False)
    assert (c == 42).all()


def test_small_cell_equals_supercell(l2k2):
    """Energy per atom of a small periodic cell (thickness < rc, many images) equals that of
    its 3x3x3 replica."""
    o = Oracle(l2k2)
    rng = np.random.default_rng(8)
    cell = st.triclinic_cell(4, 0.07, rng)
    frac = np.array([[0.1, 0.2, 0.3], [0.6, 0.1, 0.4], [0.35, 0.7, 0.8], [0.8, 0.55, 0.05]])
    pos = (frac @ cell).T
    types = np.array([0, 1, 1, 0], dtype=np.int32)
    e, f, _ = o.compute(cell, types, pos, brute=True)
    reps = [(i, j, k) for i in range(3) for j in range(3) for k in range(3)]
    big = np.concatenate([pos + (np.array(r) @ cell)[:, None] for r in reps], axis=1)
    bt = np.tile(types, len(reps))
    e2, f2, _ = o.compute(cell * 3, bt, big, brute=False)
    assert np.abs(e2.reshape(len(reps), 4) - e[None, :]).max() < 1e-10
    assert np.abs(f2.reshape(3, len(reps), 4) - f[:, None, :]).max() < 1e-9
