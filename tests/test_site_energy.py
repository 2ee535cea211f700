"""Closed-form pins of the site-energy assembly E_i = shift[t_i] + sigma[t_i] +
sigma[t_i] sum_s w_s M_i[map[s]] (src/force/sus2_sh.cu:4823-4832), for the oracle (CPU) and the
engine (GPU, FP64 and FP32).

1. An atom without neighbours inside rc has E = shift[t] + sigma[t] exactly.
2. Body order 2 only (l_max = 0: every scalar is a basic moment (k, 0, 0)): for an isolated
   dimer of two different species,
     E_i = shift[t_i] + sigma[t_i] + sigma[t_i] sum_s w_s R_{k_s}(r) Y_00,
   where the expected value is built here from the file's coefficients with the closed forms
   T_n(xi) = cos(n arccos xi), xi = tanh(a (r - s) / 2), Y_00 = 1 / (2 sqrt(pi)),
   R_mu = (r - rc)^2 sum_n scaling w[t_i] w[t_j] c_{mu n} T_n (src/force/sus2_sh.cu:2484-2517,
   3171-3220).  A dropped +sigma, a wrong species index (t_j for t_i) or a wrong pair index
   fails it: the two species carry different shift, sigma and type coefficients, and the
   pair scale / (a, s) differ by ordered pair.
"""
import math

import numpy as np
import pytest

import sus2sh
from oracle.binding import Oracle

Y00 = 1.0 / (2.0 * math.sqrt(math.pi))


def _body2_spec():
    spec = sus2sh.synthetic_model(0, 3, 2, species_count=2, rb_size=6, max_dist=5.0, seed=77,
                                  scaling=0.7, scal_mode="random")
    spec.shift_coeffs = np.array([0.31, -0.47])
    spec.species_coeffs = np.array([1.3, 0.6])
    spec.type_coeffs = np.array([1.1, 0.8])
    return spec


def _expected_dimer(spec, r, ti, tj):
    S = spec.species_count
    Mr = spec.radial_funcs_count
    rc = spec.max_dist
    g = spec.graph
    scal = spec.scal_coeffs.reshape(Mr, 2, S * S)
    pair = S * ti + tj
    e = 0.0
    for s in range(g.scalar_count):
        node = g.scalars[s]
        k, l, m = g.basic_klm[node]
        assert l == 0 and m == 0 and g.scalar_body[s] == 2
        mu = k  # mu = k (L + 1) + l with L = 0
        a, sh = scal[mu, 0, pair], scal[mu, 1, pair]
        xi = math.tanh(0.5 * a * (r - sh))
        T = np.cos(np.arange(spec.rb_size) * math.acos(xi))
        c = spec.scaling * spec.type_coeffs[ti] * spec.type_coeffs[tj] * spec.radial_coeffs[mu]
        R = (r - rc) ** 2 * float(c @ T)
        e += spec.moment_coeffs[s] * R * Y00
    return spec.shift_coeffs[ti] + spec.species_coeffs[ti] * (1.0 + e)


def _dimer_structure(r):
    cell = np.diag([60.0, 60.0, 60.0])
    pos = np.array([[10.0, 10.0 + r, 40.0],   # x
                    [10.0, 10.0, 40.0],       # y
                    [10.0, 10.0, 40.0]])      # z: atom 2 is 30 A away from both
    types = np.array([0, 1, 1], dtype=np.int32)
    return pos, types, cell


def _expected_all(spec, r):
    return np.array([_expected_dimer(spec, r, 0, 1), _expected_dimer(spec, r, 1, 0),
                     spec.shift_coeffs[1] + spec.species_coeffs[1]])


@pytest.fixture(scope="module")
def body2_model(model_dir):
    spec = _body2_spec()
    assert spec.graph.product_count == 0 and spec.graph.scalar_count == 3
    return spec, sus2sh.write_mtp(spec, str(model_dir / "body2_dimer.mtp"))


@pytest.mark.parametrize("r", [2.3, 3.7, 4.9])
def test_oracle_site_energy_closed_form(body2_model, r):
    spec, path = body2_model
    pos, types, cell = _dimer_structure(r)
    e, f, v = Oracle(path).compute(cell, types, pos, pbc=(0, 0, 0), brute=True)
    want = _expected_all(spec, r)
    print(f"r={r}: oracle {e}, closed form {want}")
    assert e[2] == spec.shift_coeffs[1] + spec.species_coeffs[1]  # no neighbour: exact
    assert np.abs(e - want).max() <= 1e-13
    assert abs(e[0] - e[1]) > 1e-3  # the two sites really differ (species-indexed terms)
    assert np.abs(f[:, 2]).max() == 0.0


def test_oracle_isolated_atoms_of_a_deep_model(make_model):
    """Body-5 l3k3 model: atoms further apart than rc keep E = shift + sigma exactly."""
    path = make_model("iso_l3k3", l_max=3, k_max=3, body_order=5, rb_size=8, max_dist=5.0,
                      seed=31, scaling=0.5)
    spec = sus2sh.synthetic_model(3, 3, 5, rb_size=8, max_dist=5.0, seed=31, scaling=0.5)
    cell = np.diag([40.0, 40.0, 40.0])
    pos = np.array([[1.0, 7.0, 13.0, 30.0], [1.0, 1.0, 1.0, 30.0], [1.0, 1.0, 1.0, 30.0]])
    types = np.array([0, 1, 0, 1], dtype=np.int32)
    e, f, _ = Oracle(path).compute(cell, types, pos, pbc=(0, 0, 0), brute=True)
    want = spec.shift_coeffs[types] + spec.species_coeffs[types]
    assert np.array_equal(e, want)
    assert np.abs(f).max() == 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("fp32", [False, True])
def test_engine_site_energy_closed_form(body2_model, fp32):
    spec, path = body2_model
    for r in (2.3, 3.7, 4.9):
        pos, types, cell = _dimer_structure(r)
        print(f"[case] body-2 dimer r={r} fp32={fp32}", flush=True)
        e, f, v = sus2sh.evaluate(path, cell, types, pos, options=["A", "B", f"sus2_float={int(fp32)}"],
                                  pbc=(0, 0, 0))
        want = _expected_all(spec, r)
        print(f"  engine {e}, closed form {want}", flush=True)
        assert abs(e[2] - want[2]) <= 1e-15
        assert np.abs(e - want).max() <= (2e-6 if fp32 else 1e-12)
