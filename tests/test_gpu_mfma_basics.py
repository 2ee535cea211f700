"""GPU parity of the matrix-core basic moments (sus2_sh_basic_mfma, FP32 engine).

Per atom C[n][lm] = sum_j sigma_p (r - rc)^2 T_n(xi_p) N_l S_lm(r_hat) is one dense product
over the edges (16x16x4 FP32 MFMAs, or 16x16x16 BF16 MFMAs with FP32 accumulation) and the
basic moments are B[mu][lm] = sum_n c[mu][n] C[n][lm] (16x16x4 FP32 MFMAs): the same moments
as src/force/sus2_sh.cu:3486-3550 with R_mu = sigma_p (r - rc)^2 sum_n c[mu][n] T_n.

FP32 operands: the FP32 parity tolerances of test_gpu_parity (2e-5 eV/atom + 2e-6 |E|,
1e-4 eV/A + 2e-5 |F|max, 1e-4 eV + 2e-5 |W|max).  BF16 operands (8-bit significands, relative
rounding up to 2^-9 per operand) carry the stated BF16 bound 3e-2 (|X|max + 1) for X = E, F, W:
the moments of odd l nearly cancel over the neighbours of a jittered fcc site, so their BF16
relative error is far above 2^-9 (measured: 1.3 % of |E|max, 0.9 % of |F|max on the l3k3
rb 8 case).  Each BF16 case also checks that its error exceeds the FP32 operands' error: the
BF16 kernel did run.
Cases cover radial rows of 8, 12 and 16 coefficients, both Chebyshev maps (tanh and the
rational map), l_max 1-3, atoms with more than 64 neighbours (several 64-edge rounds) and an
atom with no neighbour.
"""
import numpy as np
import pytest

import sus2sh
from sus2sh import structures as st
from oracle.binding import Oracle

pytestmark = pytest.mark.gpu

CASES = {
    "l3k3_rb8": dict(model=dict(l_max=3, k_max=3, body_order=5, rb_size=8, max_dist=5.0,
                                seed=12, scaling=0.5), struct=(6, 3.85, 0.08)),
    "l2k2_rb8": dict(model=dict(l_max=2, k_max=2, body_order=5, rb_size=8, max_dist=5.0,
                                seed=11), struct=(5, 3.80, 0.05)),
    "l3k2_rb12_rational": dict(model=dict(l_max=3, k_max=2, body_order=4, rb_size=12,
                                          max_dist=5.5, seed=21, scaling=0.4,
                                          radial_basis_type="RBChebyshev_sss_rational"),
                               struct=(5, 3.85, 0.08)),
    "l1k4_rb14": dict(model=dict(l_max=1, k_max=4, body_order=4, rb_size=14, max_dist=5.0,
                                 seed=22, scaling=0.5), struct=(5, 3.85, 0.08)),
    # rc 7.5 A: ~134 neighbours per atom, three 64-edge rounds per atom
    "l3k3_rb10_rc75": dict(model=dict(l_max=3, k_max=3, body_order=5, rb_size=10,
                                      max_dist=7.5, seed=23, scaling=0.12),
                           struct=(5, 3.85, 0.08)),
}


def _run(path, cell, types, pos, opts):
    print(f"[case] model={path} N={len(types)} options={opts}", flush=True)
    return sus2sh.evaluate(path, cell, types, pos, options=opts)


def _errors(ref, got):
    e0, f0, v0 = ref
    e1, f1, v1 = got
    de, df, dv = np.abs(e1 - e0).max(), np.abs(f1 - f0).max(), np.abs(v1 - v0).max()
    print(f"  max|dE|={de:.3e} max|dF|={df:.3e} max|dW|={dv:.3e} "
          f"(|E|max={np.abs(e0).max():.3f} |F|max={np.abs(f0).max():.3f})", flush=True)
    return de, df, dv


@pytest.mark.parametrize("mode", ["fp32", "bf16"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_mfma_basics_match_oracle(make_model, name, mode):
    case = CASES[name]
    path = make_model("mfma_" + name, **case["model"])
    n, a, sigma = case["struct"]
    pos, types, cell = st.fcc_alloy(n, n, n, a, [0.5, 0.5], sigma, 41, 42)
    ref = Oracle(path).compute(cell, types, pos)
    opts = ["Cu", "Zr", "sus2_float=1", f"sus2_sh_basic_mfma={mode}"]
    de, df, dv = _errors(ref, _run(path, cell, types, pos, opts))
    e0, f0, v0 = ref
    if mode == "fp32":
        assert de <= 2e-5 + 2e-6 * np.abs(e0).max()
        assert df <= 1e-4 + 2e-5 * np.abs(f0).max()
        assert dv <= 1e-4 + 2e-5 * np.abs(v0).max()
    else:
        assert de <= 3e-2 * (np.abs(e0).max() + 1)
        assert df <= 3e-2 * (np.abs(f0).max() + 1)
        assert dv <= 3e-2 * (np.abs(v0).max() + 1)
        opts32 = opts[:-1] + ["sus2_sh_basic_mfma=fp32"]
        assert df > _errors(ref, _run(path, cell, types, pos, opts32))[1]


def test_mfma_basics_isolated_atom(make_model):
    """A dimer plus an atom with no neighbour inside rc (zero-edge atom: its moments are 0
    and its energy is shift + sigma) in an open box."""
    path = make_model("mfma_l3k3_rb8", **CASES["l3k3_rb8"]["model"])
    pos = np.array([[0.0, 2.6, 30.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    types = np.array([0, 1, 1], dtype=np.int32)
    cell = np.diag([60.0, 60.0, 60.0])
    ref = Oracle(path).compute(cell, types, pos, pbc=(0, 0, 0))
    got = sus2sh.evaluate(path, cell, types, pos, options=["Cu", "Zr", "sus2_sh_basic_mfma=1"],
                          pbc=(0, 0, 0))
    de, df, dv = _errors(ref, got)
    assert de <= 2e-5 + 2e-6 * np.abs(ref[0]).max() and df <= 1e-4 + 2e-5 * np.abs(ref[1]).max()


def test_mfma_basics_rejected_where_unsupported(make_model):
    """FP64 engines, non-uniform radial scalings and l_max > 3 refuse the option cleanly."""
    p3 = make_model("mfma_l3k3_rb8", **CASES["l3k3_rb8"]["model"])
    pos, types, cell = st.fcc_alloy(3, 3, 3, 3.85, [0.5, 0.5], 0.05, 1, 2)
    bad = [(p3, ["Cu", "Zr", "sus2_float=0", "sus2_sh_basic_mfma=1"]),
           (make_model("mfma_nonuniform", l_max=3, k_max=2, body_order=4, rb_size=6,
                       max_dist=4.6, seed=14, scal_mode="random"), ["Cu", "Zr", "sus2_sh_basic_mfma=1"]),
           (make_model("mfma_l4", l_max=4, k_max=2, body_order=4, rb_size=8, max_dist=5.0,
                       seed=15), ["Cu", "Zr", "sus2_sh_basic_mfma=2"])]
    for path, opts in bad:
        with pytest.raises(sus2sh.Sus2shError) as err:
            sus2sh.evaluate(path, cell, types, pos, options=opts)
        assert "sus2_sh_basic_mfma" in err.value.message


def test_autotuner_times_the_mfma_basics(make_model):
    """From 512 atoms the autotuner also times the FP32 matrix-core kernel and keeps it when it
    is the fastest basic-moment kernel; either way the result matches the oracle."""
    from sus2sh.hipmem import DeviceArray, synchronize
    path = make_model("mfma_l3k3_rb8", **CASES["l3k3_rb8"]["model"])
    pos, types, cell = st.fcc_alloy(8, 8, 8, 3.85, [0.5, 0.5], 0.08, 43, 44)
    ref = Oracle(path).compute(cell, types, pos)
    n = len(types)
    pot = sus2sh.SUS2_SH(path, n, ["Cu", "Zr"])
    t = DeviceArray.from_numpy(np.asarray(types, dtype=np.int32))
    x = DeviceArray.from_numpy(np.ascontiguousarray(pos, dtype=np.float64))
    e = DeviceArray.zeros((n,), np.float64)
    f = DeviceArray.zeros((3, n), np.float64)
    v = DeviceArray.zeros((9, n), np.float64)
    pot.compute(sus2sh.box_struct(cell), t, x, e, f, v)
    synchronize()
    shape, times = pot.edge_shapes()
    pot.close()
    print(f"  kept {shape}; mfma fp32 {times.get(('basic_mfma_fp32',))} ms, vector 256x1 "
          f"{times.get(('basic', 256, 1))} ms", flush=True)
    assert ("basic_mfma_fp32",) in times
    best_vec = min(ms for k, ms in times.items() if k[0] == "basic")
    assert shape["basic_mfma"] == (1 if times[("basic_mfma_fp32",)] < best_vec else 0)
    de, df, dv = _errors(ref, (e.numpy(), f.numpy(), v.numpy()))
    assert de <= 2e-5 + 2e-6 * np.abs(ref[0]).max() and df <= 1e-4 + 2e-5 * np.abs(ref[1]).max()
