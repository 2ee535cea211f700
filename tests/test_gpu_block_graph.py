"""GPU parity of the block graph kernel (sus2_sh_graph_variant=4) through the C ABI.

The kernel evaluates the depth <= 2 product graph per tensor: layer-1 tensors as real
Clebsch-Gordan blocks of basic tensors (compile-time couplings), then per quadratic-form tensor
its adjoint K' v, its energy share and its reverse pass, the basic adjoints summed with LDS
atomics (semantics src/force/sus2_sh.cu:4799-4846).  Every shape (4, 8, 16 waves per 64-atom
tile) is pinned with sus2_sh_block_waves and compared with the FP64 oracle in both precisions;
a 4,096-atom case runs 64 tiles and the autotuner (which from 4,096 atoms times the block
shapes beside the lane and row kernels).

Tolerances as in test_gpu_parity: FP64 1e-9 eV/atom, 1e-8 eV/A, 1e-8 eV;
FP32 2e-5 eV/atom + 2e-6 |E|, 1e-4 eV/A + 2e-5 |F|max, 1e-4 eV + 2e-5 |W|max.
"""
import numpy as np
import pytest

import sus2sh
from sus2sh import structures as st
from oracle.binding import Oracle

pytestmark = pytest.mark.gpu

CASES = {
    "l3k3_b5": dict(model=dict(l_max=3, k_max=3, body_order=5, rb_size=8, max_dist=5.0,
                               seed=12, scaling=0.5), struct=(5, 3.85, 0.08)),
    "l2k2_b5": dict(model=dict(l_max=2, k_max=2, body_order=5, rb_size=8, max_dist=5.0,
                               seed=11), struct=(4, 3.80, 0.05)),
    "l3k2_b4_nonuniform": dict(model=dict(l_max=3, k_max=2, body_order=4, rb_size=6,
                                          max_dist=4.6, seed=14, scal_mode="random"),
                               struct=(4, 3.7, 0.1)),
    "l3k3_b5_3species_rational": dict(model=dict(l_max=3, k_max=3, body_order=5, rb_size=10,
                                                 max_dist=5.5, seed=21, species_count=3,
                                                 radial_basis_type="RBChebyshev_sss_rational",
                                                 scaling=0.4), struct=(4, 3.85, 0.08)),
    # 128 basics: the basic adjoints alone take 64 KiB of LDS, too large a tile in both
    # precisions (the engine refuses the variant cleanly)
    "l3k8_b4_chunks": dict(model=dict(l_max=3, k_max=8, body_order=4, rb_size=8, max_dist=5.0,
                                      seed=19, scaling=0.4), struct=(4, 3.85, 0.08)),
    "l1k4_b3": dict(model=dict(l_max=1, k_max=4, body_order=3, rb_size=6, max_dist=4.6,
                               seed=22), struct=(4, 3.7, 0.08)),
}


def _compare(ref, got, fp32):
    e0, f0, v0 = ref
    e1, f1, v1 = got
    de, df, dv = np.abs(e1 - e0).max(), np.abs(f1 - f0).max(), np.abs(v1 - v0).max()
    print(f"  max|dE|={de:.3e} max|dF|={df:.3e} max|dW|={dv:.3e} "
          f"(|E|max={np.abs(e0).max():.3f} |F|max={np.abs(f0).max():.3f})", flush=True)
    if fp32:
        assert de <= 2e-5 + 2e-6 * np.abs(e0).max()
        assert df <= 1e-4 + 2e-5 * np.abs(f0).max()
        assert dv <= 1e-4 + 2e-5 * np.abs(v0).max()
    else:
        assert de <= 1e-9 and df <= 1e-8 and dv <= 1e-8


def _evaluate(path, cell, types, pos, opts):
    from sus2sh.hipmem import DeviceArray, synchronize
    n = len(types)
    print(f"[case] model={path} N={n} options={opts}", flush=True)
    pot = sus2sh.SUS2_SH(path, n, opts)
    t = DeviceArray.from_numpy(np.asarray(types, dtype=np.int32))
    x = DeviceArray.from_numpy(np.ascontiguousarray(pos, dtype=np.float64))
    e = DeviceArray.zeros((n,), np.float64)
    f = DeviceArray.zeros((3, n), np.float64)
    v = DeviceArray.zeros((9, n), np.float64)
    pot.compute(sus2sh.box_struct(cell), t, x, e, f, v)
    synchronize()
    info = (pot.info()["graph_variant"], pot.tuning())
    out = (e.numpy(), f.numpy(), v.numpy())
    pot.close()
    return out, info


@pytest.mark.parametrize("fp32", [True, False])
@pytest.mark.parametrize("name", sorted(CASES))
def test_block_kernel_matches_oracle(make_model, name, fp32):
    case = CASES[name]
    path = make_model("blk_" + name, **case["model"])
    # the tile's LDS: B basic adjoints in double plus the program's Real slots, <= 160 KiB
    rb = 4 if fp32 else 8
    info = sus2sh.block_check(path, real_bytes=rb)[0]
    B = sus2sh.model_check(path)["basic_count"]
    fits = info["ok"] == 1 and B * 64 * 8 + max(info["slots"], B + 16) * 64 * rb <= 160 * 1024
    assert fits == (name != "l3k8_b4_chunks")
    n, a, sigma = case["struct"]
    S = case["model"].get("species_count", 2)
    pos, types, cell = st.fcc_alloy(n, n, n, a, [1.0 / S] * S, sigma, 51, 52)
    opts = ["Cu", "Zr", "Ni"][:S] + [f"sus2_float={int(fp32)}", "sus2_sh_graph_variant=4"]
    if not fits:
        with pytest.raises(sus2sh.Sus2shError) as err:
            sus2sh.evaluate(path, cell, types, pos, options=opts)
        assert "no block program" in err.value.message
        return
    ref = Oracle(path).compute(cell, types, pos)
    got, (variant, _) = _evaluate(path, cell, types, pos, opts)
    assert variant == 4
    _compare(ref, got, fp32)


@pytest.mark.parametrize("fp32", [True, False])
@pytest.mark.parametrize("waves", [4, 8, 16])
def test_block_shapes_4096_atoms(make_model, waves, fp32):
    """64 tiles of 64 atoms (the last one full) with every shape pinned."""
    path = make_model("blk_l3k3_b5", **CASES["l3k3_b5"]["model"])
    pos, types, cell = st.fcc_alloy(16, 16, 4, 3.85, [0.5, 0.5], 0.08, 53, 54)
    assert len(types) == 4096
    ref = Oracle(path).compute(cell, types, pos)
    opts = ["Cu", "Zr", f"sus2_float={int(fp32)}", "sus2_sh_graph_variant=4",
            f"sus2_sh_block_waves={waves}"]
    got, (variant, tuning) = _evaluate(path, cell, types, pos, opts)
    assert variant == 4 and tuning == {}
    _compare(ref, got, fp32)


@pytest.mark.parametrize("fp32", [True, False])
def test_autotuner_times_block_shapes(make_model, fp32):
    """From 4,096 atoms the autotuner times the block shapes (keys (4, c)) with the other
    graph kernels; whichever it keeps matches the oracle.  4,500 atoms: a partial last tile."""
    path = make_model("blk_l3k3_b5", **CASES["l3k3_b5"]["model"])
    pos, types, cell = st.fcc_alloy(15, 15, 5, 3.85, [0.5, 0.5], 0.08, 55, 56)
    assert len(types) == 4500
    ref = Oracle(path).compute(cell, types, pos)
    got, (variant, tuning) = _evaluate(path, cell, types, pos, ["Cu", "Zr", f"sus2_float={int(fp32)}"])
    print(f"  kept variant {variant}; block {[(k, v) for k, v in tuning.items() if k[0] == 4]}",
          flush=True)
    assert all((4, c) in tuning for c in range(3))
    best = min(tuning, key=tuning.get)
    assert variant == best[0]
    _compare(ref, got, fp32)


def test_block_variant_rejections(make_model):
    """No block program (l_max 4), or deterministic mode: the option is refused cleanly."""
    pos, types, cell = st.fcc_alloy(3, 3, 3, 3.85, [0.5, 0.5], 0.05, 1, 2)
    p4 = make_model("blk_l4k2", l_max=4, k_max=2, body_order=4, rb_size=6, max_dist=4.6, seed=15)
    p3 = make_model("blk_l3k3_b5", **CASES["l3k3_b5"]["model"])
    for path, opts in [(p4, ["Cu", "Zr", "sus2_sh_graph_variant=4"]),
                       (p3, ["Cu", "Zr", "sus2_sh_graph_variant=4", "sus2_sh_deterministic=1"])]:
        with pytest.raises(sus2sh.Sus2shError) as err:
            sus2sh.evaluate(path, cell, types, pos, options=opts)
        assert "sus2_sh_graph_variant=4" in err.value.message


@pytest.mark.parametrize("fp32", [True, False])
def test_block_program_follows_set_params(model_dir, fp32):
    """The block program folds the energy weights into its records: after set_params a plain
    evaluation (no dE/dtheta) must use the new weights -- checked against the oracle on a model
    file written with the same parameters."""
    from sus2sh.hipmem import DeviceArray, synchronize
    spec = sus2sh.synthetic_model(3, 3, 5, rb_size=8, max_dist=5.0, seed=12, scaling=0.5)
    path = sus2sh.write_mtp(spec, str(model_dir / "blk_setp_a.mtp"))
    theta = sus2sh.model_gen.param_vector(spec)
    rng = np.random.default_rng(5)
    theta2 = theta * (1.0 + 0.2 * rng.standard_normal(len(theta)))
    path2 = sus2sh.write_mtp(sus2sh.model_gen.with_params(spec, theta2),
                             str(model_dir / "blk_setp_b.mtp"))
    pos, types, cell = st.fcc_alloy(4, 4, 4, 3.85, [0.5, 0.5], 0.08, 57, 58)
    n = len(types)
    ref = Oracle(path2).compute(cell, types, pos)
    print(f"[case] set_params then compute, fp32={fp32} N={n}", flush=True)
    pot = sus2sh.SUS2_SH(path, n, ["Cu", "Zr", f"sus2_float={int(fp32)}", "sus2_sh_graph_variant=4"])
    t = DeviceArray.from_numpy(np.asarray(types, dtype=np.int32))
    x = DeviceArray.from_numpy(np.ascontiguousarray(pos, dtype=np.float64))
    out = []
    for th in (theta, theta2):
        pot.set_params(th)
        e = DeviceArray.zeros((n,), np.float64)
        f = DeviceArray.zeros((3, n), np.float64)
        v = DeviceArray.zeros((9, n), np.float64)
        pot.compute(sus2sh.box_struct(cell), t, x, e, f, v)
        synchronize()
        out.append((e.numpy(), f.numpy(), v.numpy()))
    pot.close()
    _compare(ref, out[1], fp32)
    assert np.abs(out[1][0] - out[0][0]).max() > 1e-3  # the weights did change the energies
