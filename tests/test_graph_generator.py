"""Standard real-CG graph generator and CG algebra (host code, CPU only).

Pins:
  * graph dimensions recorded by the reference for its own models
    (docs/gpumd-sus2-sh-interface-plan.md:24-40 and :986-992):
      l3333   : A=1349, B=48, T=5475, Q=572, rows=1301, row terms 5475
      l3322   : rows 477, row terms 1416
      l4k4_4422: rows 968, row terms 2996
      l4k5_4422: rows 1730, row terms 5410
      l6 {6,6,1,1}, k_max 1: A=67, B=49, T=67, Q=14, rows=18 -- the reference loader found
                 this model's saved graph equal to the generated standard graph
                 (docs/gpumd-sus2-sh-interface-plan.md:2044, 2217-2223)
  * textbook Clebsch-Gordan values and orthogonality;
  * real CG coefficients are real and couple real harmonics correctly (Gaunt check by
    quadrature against the oracle's independent Y_lm).
"""
import itertools
import math

import numpy as np
import pytest

import sus2sh


@pytest.mark.parametrize("name,L,K,body,blmax,expect", [
    ("l3333", 3, 3, 5, (3, 3, 3, 3), dict(A=1349, B=48, T=5475, Q=572, rows=1301)),
    ("l3322", 3, 3, 5, (3, 3, 2, 2), dict(T=1416, rows=477)),
    ("l4k4_4422", 4, 4, 5, (4, 4, 2, 2), dict(T=2996, rows=968)),
    ("l4k5_4422", 4, 5, 5, (4, 4, 2, 2), dict(T=5410, rows=1730)),
    ("l6_6611", 6, 1, 5, (6, 6, 1, 1), dict(A=67, B=49, T=67, Q=14, rows=18)),
])
def test_reference_recorded_dimensions(name, L, K, body, blmax, expect):
    g = sus2sh.standard_graph(L, K, body, blmax)
    rows = len(np.unique(g.products[:, 2]))
    got = dict(A=g.node_count, B=g.basic_count, T=g.product_count, Q=g.scalar_count, rows=rows)
    for k, v in expect.items():
        assert got[k] == v, (name, k, got[k], v)


def test_graph_structure_invariants():
    """Validator invariants of scripts/validate_sus2_sh_metadata.py:95-161 restated."""
    for L, K, body in [(2, 2, 5), (3, 3, 5), (2, 3, 6), (4, 2, 4), (1, 1, 3)]:
        g = sus2sh.standard_graph(L, K, body)
        B, A = g.basic_count, g.node_count
        left, right, tgt = g.products.T
        assert (tgt >= B).all(), "no product writes into a basic"
        assert (left <= right).all(), "left <= right after the swap"
        assert ((left < A) & (right < A) & (tgt < A)).all()
        last_def = {}
        for p, t in enumerate(tgt):
            last_def[t] = p
        for p in range(g.product_count):
            for src in (left[p], right[p]):
                if src >= B:
                    assert last_def[src] < p, "topological order"
        defined = set(range(B)) | set(tgt.tolist())
        assert set(g.scalars.tolist()) <= defined
        assert len(set(g.scalars.tolist())) == g.scalar_count
        # basics: l = L..0, k = K-1..0, m = -l..l
        expected = [(k, l, m) for l in range(L, -1, -1) for k in range(K - 1, -1, -1)
                    for m in range(-l, l + 1)]
        assert [tuple(x) for x in g.basic_klm] == expected
        assert (np.abs(g.coeffs) >= 1e-12).all()


def test_body2_and_body3_scalars():
    g = sus2sh.standard_graph(2, 2, 3)
    # body 2: the l=0 basics (k = 1, 0) ; body 3: couples of equal l with k1 <= k0
    assert list(g.scalar_body[:2]) == [2, 2]
    n3 = int((g.scalar_body == 3).sum())
    assert n3 == 3 * 3  # l in {2,1,0} x (k0,k1) in {(1,1),(1,0),(0,0)}


def test_clebsch_gordan_textbook():
    cg = sus2sh.clebsch_gordan
    assert cg(1, 0, 1, 0, 2, 0) == pytest.approx(math.sqrt(2.0 / 3.0), abs=1e-14)
    assert cg(1, 1, 1, -1, 0, 0) == pytest.approx(1.0 / math.sqrt(3.0), abs=1e-14)
    assert cg(1, 0, 1, 0, 0, 0) == pytest.approx(-1.0 / math.sqrt(3.0), abs=1e-14)
    assert cg(1, 1, 1, 0, 1, 1) == pytest.approx(1.0 / math.sqrt(2.0), abs=1e-14)
    assert cg(2, 2, 2, -2, 0, 0) == pytest.approx(1.0 / math.sqrt(5.0), abs=1e-14)
    assert cg(1, 1, 1, 1, 2, 2) == pytest.approx(1.0, abs=1e-14)
    assert cg(1, 0, 1, 0, 1, 0) == 0.0
    assert cg(1, 1, 1, 1, 1, 1) == 0.0  # m mismatch


@pytest.mark.parametrize("j1,j2", [(1, 1), (2, 1), (2, 2), (3, 2), (4, 3)])
def test_clebsch_gordan_orthogonality(j1, j2):
    cg = sus2sh.clebsch_gordan
    for J in range(abs(j1 - j2), j1 + j2 + 1):
        for Jp in range(abs(j1 - j2), j1 + j2 + 1):
            for M in range(-min(J, Jp), min(J, Jp) + 1):
                s = sum(cg(j1, m1, j2, M - m1, J, M) * cg(j1, m1, j2, M - m1, Jp, M)
                        for m1 in range(-j1, j1 + 1) if abs(M - m1) <= j2)
                assert s == pytest.approx(1.0 if J == Jp else 0.0, abs=1e-12)


def _lebedev_like_grid(n_theta=40, n_phi=80):
    """Gauss-Legendre in cos(theta) x uniform phi: exact for the polynomial degrees used."""
    x, w = np.polynomial.legendre.leggauss(n_theta)
    phi = np.arange(n_phi) * 2 * np.pi / n_phi
    pts, wts = [], []
    for xi, wi in zip(x, w):
        s = math.sqrt(1 - xi * xi)
        for p in phi:
            pts.append((s * math.cos(p), s * math.sin(p), xi))
            wts.append(wi * 2 * np.pi / n_phi)
    return np.array(pts), np.array(wts)


def test_real_cg_couples_real_harmonics():
    """Gaunt check: the L=0 real coupling of two real harmonics of degree l is diagonal
    and uniform in m, (Y_l (x) Y_l)_0 = c_l sum_m Y_lm Y_lm with |c_l| = 1/sqrt(2l+1), and
    the real CG matrices are orthogonal in the coupled index."""
    for l in range(0, 7):
        vals = set()
        for m1, m2 in itertools.product(range(-l, l + 1), repeat=2):
            c = sus2sh.real_cg(l, m1, l, m2, 0, 0)
            if m1 != m2:
                assert abs(c) < 1e-12
            else:
                vals.add(round(c, 12))
        assert len(vals) == 1
        assert abs(abs(vals.pop()) - 1.0 / math.sqrt(2 * l + 1)) < 1e-12
    for l1, l2 in [(1, 1), (2, 1), (2, 2), (3, 2), (5, 1), (6, 2)]:
        for L in range(abs(l1 - l2), l1 + l2 + 1):
            for Lp in range(abs(l1 - l2), l1 + l2 + 1):
                for M in range(-L, L + 1):
                    for Mp in range(-Lp, Lp + 1):
                        s = sum(sus2sh.real_cg(l1, a, l2, b, L, M) * sus2sh.real_cg(l1, a, l2, b, Lp, Mp)
                                for a in range(-l1, l1 + 1) for b in range(-l2, l2 + 1))
                        assert s == pytest.approx(1.0 if (L, M) == (Lp, Mp) else 0.0, abs=1e-12)


def test_real_cg_gaunt_quadrature():
    """sum_{m1 m2} C^{LM}_{m1 m2} Y_{l1 m1}(u) Y_{l2 m2}(u) is proportional to Y_LM(u) for
    every u (with an (l1,l2,L) factor), checked with the oracle's closed-form Y_lm."""
    from oracle.binding import ylm
    pts, _ = _lebedev_like_grid(6, 7)
    for l1, l2 in [(1, 1), (2, 1), (2, 2), (3, 1), (3, 2), (3, 3), (4, 2), (5, 1)]:
        for L in range(abs(l1 - l2), min(l1 + l2, 6) + 1):
            if (l1 + l2 + L) % 2:
                continue  # odd couplings vanish pointwise for products of the same vector
            ratios = []
            for u in pts:
                Y, _ = ylm(6, u)
                for M in range(-L, L + 1):
                    c = sum(sus2sh.real_cg(l1, a, l2, b, L, M) * Y[l1 * l1 + a + l1] * Y[l2 * l2 + b + l2]
                            for a in range(-l1, l1 + 1) for b in range(-l2, l2 + 1))
                    y = Y[L * L + M + L]
                    if abs(y) > 1e-3:
                        ratios.append(c / y)
            ratios = np.array(ratios)
            assert np.ptp(ratios) < 1e-10 * max(1.0, np.abs(ratios).max()), (l1, l2, L)
