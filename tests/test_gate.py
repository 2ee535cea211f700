"""Two-layer mu gate models (docs/mu-body-order-gate.md:25-63; loader
src/force/sus2_sh.cu:2153-2266, 2412-2478; per step :9821-10256).

Loader (CPU): the reference's accepted forms (mu-body-linear-combo, mu-scalar-full with the
neighbour site and shared radial scaling) load; every rejection listed by the reference
fails with its message.

Oracle pins (CPU, FP64):
  * amplitude A = 0 gives G = 1 exactly: E/F/W equal the plain model's;
  * mu-scalar-full with W[mu][q] = c_{mu,body(q)} w_q equals mu-body-linear-combo (c, w);
  * closed form on a body-2 dimer (l_max = 0, k_max = 1): s_j = Rg(r) Y00,
    G_j = 1 + A tanh(a_{t_j} W s_j), E_i = shift_i + sigma_i (1 + w G_j R(r) Y00) -- which
    neighbour's gate and which species' additive coefficient enter;
  * forces = -dE/dx by central differences, virial = strain derivative of E, rotation
    invariance of E (both modes, tanh in its nonlinear range).
Engine (GPU): E/F/W against the oracle in FP64 (1e-9 eV/atom, 1e-8 eV/A, 1e-8 eV) and FP32
(2e-5 eV/atom + 2e-6 |E|, 1e-4 eV/A + 2e-5 |F|max, 1e-4 eV + 2e-5 |W|max).  FP64 runs the gate
weights at scale 0.6 (tanh deep in its nonlinear range); FP32 at 0.1, the generator's scale:
the FP32 force error of mu-scalar-full grows with that scale through the gate layer's reverse
pass (measured on l3k3, 256 atoms: 1.2e-6 eV/A at 0.1, 4.4e-5 at 0.3, 1.7e-3 at 0.6, while
FP64 stays at 1e-13 and FP32 with A = 0 at 1.5e-7).
"""
import math
import re

import numpy as np
import pytest

import sus2sh
from sus2sh import structures as st
from oracle.binding import Oracle

Y00 = 1.0 / (2.0 * math.sqrt(math.pi))


def _gate_spec(mode, seed=5, amplitude=0.8, L=2, K=2, body=5, rb=6, rc=4.5, S=2, scale=0.6,
               **kw):
    spec = sus2sh.synthetic_model(L, K, body, species_count=S, rb_size=rb, max_dist=rc,
                                  seed=seed, scaling=0.5, **kw)
    return spec, sus2sh.gate_model(spec, mode, seed=seed + 100, amplitude=amplitude,
                                   coeff_scale=scale)


def _write(model_dir, name, spec):
    return sus2sh.write_mtp(spec, str(model_dir / (name + ".mtp")))


# ----------------------------------------------------------------------------- loader
def _text(model_dir, mode="mu-scalar-full"):
    _, g = _gate_spec(mode)
    return open(_write(model_dir, "gate_text_" + mode, g)).read()


def _check_rejected(model_dir, text, message):
    path = str(model_dir / "gate_bad.mtp")
    with open(path, "w") as fh:
        fh.write(text)
    with pytest.raises(sus2sh.Sus2shError) as exc:
        sus2sh.model_check(path)
    assert message in str(exc.value), str(exc.value)


@pytest.mark.parametrize("mode", ["mu-scalar-full", "mu-body-linear-combo", "mu-body-order"])
def test_loader_accepts_gate_modes(model_dir, mode):
    spec_mode = "mu-scalar-full" if mode == "mu-body-order" else mode
    text = _text(model_dir, spec_mode).replace("two_layer_gate_mode = " + spec_mode,
                                               "two_layer_gate_mode = " + mode)
    path = str(model_dir / ("gate_ok_" + mode + ".mtp"))
    with open(path, "w") as fh:
        fh.write(text)
    assert sus2sh.model_check(path)["scalar_count"] > 0


REJECTIONS = [
    (lambda t: t.replace("two_layer_gate_mode = mu-scalar-full", "two_layer_gate_mode = legacy"),
     "legacy gate models are not supported here"),
    (lambda t: t.replace("two_layer_gate_mode = mu-scalar-full", "two_layer_gate_mode = other"),
     "Unsupported SUS2_SH two_layer_gate_mode in GPUMD: other"),
    (lambda t: t + "two_layer_gate_include_one_body = true\n",
     "requires two_layer_gate_include_one_body = false"),
    (lambda t: t + "two_layer_gate_scale_mode = x\n", "does not support two_layer_gate_scale_mode"),
    (lambda t: t + "two_layer_gate_bias = 0.1\n", "does not support two_layer_gate_bias"),
    (lambda t: t + "two_layer_residual_enabled = true\n", "does not support residual gate fields"),
    (lambda t: t.replace("two_layer_gate_site_mode = neighbor", "two_layer_gate_site_mode = double"),
     "currently supports two_layer_gate_site_mode = neighbor"),
    (lambda t: re.sub(r"two_layer_gate_tanh_amplitude = [^\n]*", "two_layer_gate_tanh_amplitude = 1.5", t),
     "should be finite and in [0,1]"),
    (lambda t: t.replace("two_layer_gate_radial_mode = shared-radial", "two_layer_gate_radial_mode = own"),
     "requires shared-radial gate mode"),
    (lambda t: t.replace("two_layer_gate_body_order_max = 3", "two_layer_gate_body_order_max = 4"),
     "should equal sh_k_max + 1"),
    (lambda t: re.sub(r"two_layer_gate_radial_coeff_count = \d+", "two_layer_gate_radial_coeff_count = 7", t),
     "gate radial coefficient count is inconsistent"),
    (lambda t: re.sub(r"two_layer_gate_additive_coeff_count = \d+", "two_layer_gate_additive_coeff_count = 3", t),
     "additive coefficient count is inconsistent"),
    (lambda t: re.sub(r"two_layer_gate_weight_count = \d+", "two_layer_gate_weight_count = 5", t),
     "scalar index/weight dimensions are inconsistent"),
    (lambda t: t + "two_layer_gate_body_mix_weight_count = 12\n",
     "mu-scalar-full gate should not contain body-mix weights"),
    (lambda t: re.sub(r"two_layer_gate_scalar_indices = \{(\d+)", r"two_layer_gate_scalar_indices = {9999", t),
     "gate scalar index is out of range"),
]


@pytest.mark.parametrize("k", range(len(REJECTIONS)))
def test_loader_rejections_carry_the_reference_messages(model_dir, k):
    edit, message = REJECTIONS[k]
    _check_rejected(model_dir, edit(_text(model_dir)), message)


def test_loader_requires_every_body_bucket_and_scalar_info(model_dir):
    spec, g = _gate_spec("mu-scalar-full")
    # keep only body-2 gate scalars: the body-3 bucket (k_max + 1 = 3) is missing
    idx = np.asarray(g.tail["two_layer_gate_scalar_indices"])
    keep = spec.graph.scalar_body[idx] == 2
    Mr = spec.radial_funcs_count
    g.tail["two_layer_gate_scalar_indices"] = idx[keep]
    g.tail["two_layer_gate_weights"] = np.asarray(g.tail["two_layer_gate_weights"]).reshape(Mr, -1)[:, keep].reshape(-1)
    g.tail["two_layer_gate_weight_count"] = int(Mr * keep.sum())
    text = open(_write(model_dir, "gate_bucket", g)).read()
    _check_rejected(model_dir, text, "misses a required body-order bucket")
    text = open(_write(model_dir, "gate_info", _gate_spec("mu-scalar-full")[1])).read()
    text = re.sub(r"sh_scalar_info_count = [^\n]*\n", "", text)
    _check_rejected(model_dir, text, "requires sh_scalar_info metadata")


def test_deterministic_option_rejects_gate_models(model_dir):
    """Deterministic summation covers the plain and ZBL force paths; the gate layer's dE/dh
    scatter has no fixed-point form, so a handle asking for both fails at creation (before
    any device call, so this runs without a GPU)."""
    path = _write(model_dir, "gate_det", _gate_spec("mu-scalar-full")[1])
    with pytest.raises(sus2sh.Sus2shError) as exc:
        sus2sh.SUS2_SH(path, 8, ["Cu", "Zr", "sus2_sh_deterministic=1"])
    assert "sus2_sh_deterministic=1 is not available for two-layer gate models" in str(exc.value)


# ----------------------------------------------------------------------------- oracle
def _struct(S=2, n=3, a=3.7, sigma=0.12, seed=1):
    return st.fcc_alloy(n, n, n, a, [1.0 / S] * S, sigma, seed, seed + 1)


def test_oracle_zero_amplitude_is_the_plain_model(model_dir):
    spec, g = _gate_spec("mu-scalar-full", amplitude=0.0)
    pos, types, cell = _struct()
    a = Oracle(_write(model_dir, "gate_a0", g)).compute(cell, types, pos)
    b = Oracle(_write(model_dir, "gate_plain", spec)).compute(cell, types, pos)
    for x, y in zip(a, b):
        assert np.abs(x - y).max() <= 1e-13 * max(1.0, np.abs(y).max())


def test_oracle_modes_agree(model_dir):
    spec, g = _gate_spec("mu-body-linear-combo")
    K, Mr = spec.graph.k_max, spec.radial_funcs_count
    idx = np.asarray(g.tail["two_layer_gate_scalar_indices"])
    body = spec.graph.scalar_body[idx] - 2
    mix = np.asarray(g.tail["two_layer_gate_body_mix_weights"]).reshape(Mr, K)
    w = np.asarray(g.tail["two_layer_gate_weights"])
    full = sus2sh.gate_model(spec, "mu-scalar-full", seed=0)
    full.tail.update({k: v for k, v in g.tail.items() if k.startswith("two_layer_gate_") and
                      "weight" not in k and "mode" not in k})
    full.tail["two_layer_gate_scalar_indices"] = idx
    full.tail["two_layer_gate_weights"] = (mix[:, body] * w[None, :]).reshape(-1)
    full.tail["two_layer_gate_weight_count"] = Mr * len(idx)
    pos, types, cell = _struct()
    a = Oracle(_write(model_dir, "gate_combo", g)).compute(cell, types, pos)
    b = Oracle(_write(model_dir, "gate_full_eq", full)).compute(cell, types, pos)
    for x, y in zip(a, b):
        assert np.abs(x - y).max() <= 1e-12 * max(1.0, np.abs(y).max())


def test_oracle_dimer_closed_form(model_dir):
    spec = sus2sh.synthetic_model(0, 1, 2, species_count=2, rb_size=6, max_dist=5.0, seed=41,
                                  scaling=0.8, scal_mode="random")
    spec.shift_coeffs = np.array([0.2, -0.3])
    spec.species_coeffs = np.array([1.4, 0.7])
    spec.type_coeffs = np.array([1.2, 0.9])
    spec.moment_coeffs = np.array([0.9])
    g = sus2sh.gate_model(spec, "mu-scalar-full", seed=42, amplitude=0.7, coeff_scale=1.5)
    g.tail["two_layer_gate_additive_coeffs"] = np.array([1.3, 0.6])
    path = _write(model_dir, "gate_dimer", g)
    Wg = float(np.asarray(g.tail["two_layer_gate_weights"])[0])
    gr = np.asarray(g.tail["two_layer_gate_radial_coeffs"])
    add = g.tail["two_layer_gate_additive_coeffs"]
    scal = spec.scal_coeffs.reshape(1, 2, 4)
    rc = spec.max_dist

    def radial(coef, r, ti, tj):
        pair = 2 * ti + tj
        xi = math.tanh(0.5 * scal[0, 0, pair] * (r - scal[0, 1, pair]))
        T = np.cos(np.arange(spec.rb_size) * math.acos(xi))
        return (r - rc) ** 2 * spec.scaling * spec.type_coeffs[ti] * spec.type_coeffs[tj] * float(coef @ T)

    for r in (2.4, 3.6, 4.8):
        cell = np.diag([60.0, 60.0, 60.0])
        pos = np.array([[10.0, 10.0 + r, 40.0], [10.0, 10.0, 40.0], [10.0, 10.0, 40.0]])
        types = np.array([0, 1, 1], dtype=np.int32)
        e, f, v = Oracle(path).compute(cell, types, pos, pbc=(0, 0, 0), brute=True)
        want = []
        for i, j in ((0, 1), (1, 0)):
            s_j = radial(gr, r, types[j], types[i]) * Y00          # gate moment of atom j
            G_j = 1.0 + 0.7 * math.tanh(add[types[j]] * Wg * s_j)
            B_i = G_j * radial(spec.radial_coeffs[0], r, types[i], types[j]) * Y00
            want.append(spec.shift_coeffs[types[i]] + spec.species_coeffs[types[i]] *
                        (1.0 + spec.moment_coeffs[0] * B_i))
        want.append(spec.shift_coeffs[1] + spec.species_coeffs[1])
        print(f"r={r}: {e} vs {want}")
        assert np.abs(e - np.array(want)).max() <= 1e-13
        assert r > 4.0 or abs(want[0] - (spec.shift_coeffs[0] + spec.species_coeffs[0] * (1.0 + spec.moment_coeffs[0] *
                   radial(spec.radial_coeffs[0], r, 0, 1) * Y00))) > 1e-4  # the gate is active


@pytest.mark.parametrize("mode", ["mu-scalar-full", "mu-body-linear-combo"])
def test_oracle_forces_virial_rotation(model_dir, mode):
    spec, g = _gate_spec(mode, L=3, K=2, body=5, rb=5, rc=4.4, S=2, scale=0.8, seed=9)
    o = Oracle(_write(model_dir, "gate_fd_" + mode, g))
    pos, types, cell = _struct(n=2, a=3.5, seed=3)
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
    W = v.sum(axis=1)
    order = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2), (1, 0), (2, 0), (2, 1)]
    for k, (a, b) in enumerate(order):
        def energy(eps):
            F = np.eye(3)
            F[a, b] += eps
            return o.compute(cell @ F.T, types, F @ pos)[0].sum()
        dE = (energy(1e-6) - energy(-1e-6)) / 2e-6
        assert W[k] == pytest.approx(-dE, abs=1e-6 * max(1.0, abs(dE)))
    q, _ = np.linalg.qr(np.random.default_rng(9).normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    e3, f3, _ = o.compute(cell @ q.T, types, q @ pos)
    assert np.abs(e3 - e).max() < 1e-10
    assert np.abs(f3 - q @ f).max() < 1e-9


# ----------------------------------------------------------------------------- engine
GPU_CASES = {
    "l2k2_full": dict(mode="mu-scalar-full", L=2, K=2, body=5, rb=8, rc=5.0, S=2, seed=11),
    "l2k2_combo": dict(mode="mu-body-linear-combo", L=2, K=2, body=5, rb=8, rc=5.0, S=2, seed=12),
    "l3k3_full": dict(mode="mu-scalar-full", L=3, K=3, body=5, rb=8, rc=5.0, S=2, seed=13),
    "l3k3_combo": dict(mode="mu-body-linear-combo", L=3, K=3, body=5, rb=8, rc=5.0, S=2, seed=14),
    "l4k2_rational_S3": dict(mode="mu-body-linear-combo", L=4, K=2, body=4, rb=10, rc=5.5, S=3,
                             seed=15, radial_basis_type="RBChebyshev_sss_rational"),
}


def _compare(ref, got, fp32):
    (e0, f0, v0), (e1, f1, v1) = ref, got
    de, df, dv = np.abs(e1 - e0).max(), np.abs(f1 - f0).max(), np.abs(v1 - v0).max()
    print(f"  max|dE|={de:.3e} max|dF|={df:.3e} max|dW|={dv:.3e} (|F|max={np.abs(f0).max():.3f})",
          flush=True)
    if fp32:
        assert de <= 2e-5 + 2e-6 * np.abs(e0).max()
        assert df <= 1e-4 + 2e-5 * np.abs(f0).max()
        assert dv <= 1e-4 + 2e-5 * np.abs(v0).max()
    else:
        assert de <= 1e-9 and df <= 1e-8 and dv <= 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("name", sorted(GPU_CASES))
def test_engine_gate_matches_oracle(model_dir, name, fp32):
    kw = dict(GPU_CASES[name])
    mode, S = kw.pop("mode"), kw["S"]
    _, g = _gate_spec(mode, scale=0.1 if fp32 else 0.6, **kw)
    path = _write(model_dir, f"gate_gpu_{name}_{int(fp32)}", g)
    pos, types, cell = st.fcc_alloy(4, 4, 4, 3.8, [1.0 / S] * S, 0.1, 7, 8)
    ref = Oracle(path).compute(cell, types, pos)
    print(f"[case] gate {name} fp32={fp32} N={len(types)}", flush=True)
    got = sus2sh.evaluate(path, cell, types, pos,
                          options=[f"T{s}" for s in range(S)] + [f"sus2_float={int(fp32)}"])
    _compare(ref, got, fp32)


@pytest.mark.gpu
@pytest.mark.parametrize("fp32", [False, True])
def test_engine_gate_small_cells(model_dir, fp32):
    """Cells thinner than 2 rc: the gate multiplier of a neighbour image is its atom's."""
    from test_gpu_parity import _separated_positions
    _, g = _gate_spec("mu-scalar-full", L=3, K=3, body=5, rb=8, rc=5.0, seed=21,
                      scale=0.1 if fp32 else 0.6)
    path = _write(model_dir, f"gate_small_{int(fp32)}", g)
    orc = Oracle(path)
    rng = np.random.default_rng(5)
    for n in (8, 23):
        cell = st.triclinic_cell(n, 0.065, rng)
        pos = _separated_positions(cell, n, 1.8, rng)
        types = rng.integers(0, 2, size=n).astype(np.int32)
        ref = orc.compute(cell, types, pos, brute=True)
        print(f"[case] gate small cell n={n} fp32={fp32}", flush=True)
        got = sus2sh.evaluate(path, cell, types, pos, options=["A", "B", f"sus2_float={int(fp32)}"])
        _compare(ref, got, fp32)
