"""Lane program -- the wave-uniform instruction stream of the one-atom-per-lane graph kernel
(graph variant 3) -- executed on the host through the C ABI (sus2sh_lane_check) against the
FP64 oracle's flat product graph (oracle_graph: src/force/sus2_sh.cu:4799-4846) on seeded
basic vectors.

Bar: E' = sum_s w_s M[map[s]] and dE'/dbasic agree to 1e-11 relative to their scale (both
in double; the lane program sums in another order: components of the layer-2 quadratic form
and E' = 1/2 sum_n v_n (g_n + d_n)).  Three-layer (body-order 6) graphs add the layer-2 tensor
nodes Y that feed layer-3 products to the quadratic form, with their layer-1 operands P kept in
persistent slots.
"""
import dataclasses

import numpy as np
import pytest

import sus2sh
from oracle.binding import Oracle

STANDARD = [
    # (l_max, k_max, body_order, body_l_max)
    (0, 1, 2, None),
    (0, 3, 3, None),
    (1, 1, 3, None),
    (1, 2, 4, None),
    (2, 2, 5, None),
    (2, 3, 4, None),
    (3, 3, 5, (3, 3, 3, 3)),
    (3, 2, 5, (3, 2, 1, 0)),
    (4, 2, 4, None),
    (4, 4, 5, (4, 4, 2, 2)),
]


def _check(path, max_slots=1024, n=5, seed=0, expect_ok=True):
    """Every wave split (1, 2, 4, 8 waves per tile) of the program against the oracle."""
    orc = Oracle(path)
    B = orc.dims["B"]
    rng = np.random.default_rng(seed)
    basics = rng.normal(0.0, 0.4, size=(n, B))
    e0, g0 = orc.graph(basics)
    scale = max(1.0, np.abs(e0).max(), np.abs(g0).max())
    for waves in (1, 2, 4, 8):
        info, e, g = sus2sh.lane_check(path, basics, max_slots=max_slots, waves=waves)
        print(f"[lane] {path} waves={waves}: {info}", flush=True)
        if not expect_ok:
            assert info["ok"] == 0
            continue
        assert info["ok"] == 1
        assert np.abs(e - e0).max() <= 1e-11 * scale
        assert np.abs(g - g0).max() <= 1e-11 * scale
    return info


@pytest.mark.parametrize("L,K,body,blm", STANDARD)
def test_standard_graphs(make_model, L, K, body, blm):
    path = make_model(f"lane_l{L}k{K}b{body}_{blm}", l_max=L, k_max=K, body_order=body,
                      body_l_max=blm, rb_size=4, max_dist=4.0, seed=100 + 7 * L + K + body,
                      species_count=2)
    info = _check(path)
    assert info["instructions"] > 0


def test_l3333_structure(make_model):
    """The l3333 graph (C3 workload) splits into 64 (L, M) components of at most 33 nodes and
    fits the 160-slot budget (40 KiB of LDS per 64-atom FP32 workgroup)."""
    path = make_model("lane_l3333", l_max=3, k_max=3, body_order=5, body_l_max=(3, 3, 3, 3),
                      rb_size=8, max_dist=5.0, seed=3003, scaling=0.5)
    info = _check(path, max_slots=160)
    assert info["components"] == 64
    assert info["largest_component"] == 33
    assert info["slots"] <= 160


# every lane shape the engine's autotuner tries (waves, accumulators, slot budget, atoms per
# lane)
LANE_SHAPES = [(4, 4, 160, 1), (8, 2, 160, 1), (4, 2, 160, 1), (8, 2, 192, 1), (8, 2, 256, 1),
               (2, 4, 160, 1), (8, 2, 208, 1), (8, 2, 160, 2), (4, 2, 160, 2), (8, 2, 208, 2),
               (16, 2, 192, 1), (16, 2, 320, 1)]


@pytest.mark.parametrize("waves,nslot,slots,atoms", LANE_SHAPES)
def test_l3333_every_tuned_shape_and_encoding(make_model, waves, nslot, slots, atoms):
    """The l3333 program at each autotuned shape, run three ways on the host against the
    oracle's flat graph: the compiled program and its FP64 device encoding (1e-11 of the
    scale), and its FP32 device encoding -- the records the benchmarked FP32 kernel reads,
    coefficients rounded to float -- to 2e-6 of the scale (float rounding of ~10^4
    coefficients)."""
    path = make_model("lane_l3333", l_max=3, k_max=3, body_order=5, body_l_max=(3, 3, 3, 3),
                      rb_size=8, max_dist=5.0, seed=3003, scaling=0.5)
    orc = Oracle(path)
    B = orc.dims["B"]
    rng = np.random.default_rng(waves * 1000 + nslot * 100 + slots + atoms)
    basics = rng.normal(0.0, 0.4, size=(4, B))
    e0, g0 = orc.graph(basics)
    scale = max(1.0, np.abs(e0).max(), np.abs(g0).max())
    for real_bytes, tol in ((0, 1e-11), (8, 1e-11), (4, 2e-6)):
        info, e, g = sus2sh.lane_check(path, basics, max_slots=slots, waves=waves, nslot=nslot,
                                       real_bytes=real_bytes, atoms=atoms)
        print(f"[lane] l3333 shape=({waves},{nslot},{slots},{atoms}) bytes={real_bytes}: {info}",
              flush=True)
        assert info["ok"] == 1 and info["slots"] <= slots
        assert np.abs(e - e0).max() <= tol * scale
        assert np.abs(g - g0).max() <= tol * scale


BODY6 = [
    # (l_max, k_max, body_l_max, slot budget): full-l graphs (many Y / P nodes) and the
    # production shape with a last intermediate of l <= 1 ({.., 1}: 73 Y, 45 P nodes)
    (2, 2, None, 1024),
    (1, 3, None, 1024),
    (3, 3, (3, 3, 3, 3, 1), 384),
    (4, 3, (4, 4, 4, 4, 1), 384),
]


@pytest.mark.parametrize("L,K,blm,slots", BODY6)
def test_body6_graphs(make_model, L, K, blm, slots):
    """Three-layer graphs: every wave split against the oracle's flat graph, and the FP32
    device encoding of the 8-wave program at 2e-6 of the scale."""
    path = make_model(f"lane_body6_l{L}k{K}_{blm}", l_max=L, k_max=K, body_order=6,
                      body_l_max=blm, rb_size=4, max_dist=4.0, seed=5 + L + K, species_count=2)
    assert sus2sh.model_check(path)["layer_count"] == 3
    info = _check(path, max_slots=slots, n=3)
    assert info["instructions"] > 0
    orc = Oracle(path)
    basics = np.random.default_rng(11).normal(0.0, 0.4, size=(3, orc.dims["B"]))
    e0, g0 = orc.graph(basics)
    scale = max(1.0, np.abs(e0).max(), np.abs(g0).max())
    info, e, g = sus2sh.lane_check(path, basics, max_slots=slots, waves=8, nslot=2, real_bytes=4)
    assert info["ok"] == 1
    assert np.abs(e - e0).max() <= 2e-6 * scale
    assert np.abs(g - g0).max() <= 2e-6 * scale


def test_non_standard_depth3_graph(model_dir):
    """A three-layer graph that is not the standard one: perturbed coefficients, scalars
    mapped to random nodes of every layer (so some layer-2 operands of layer-3 products are
    mapped themselves, and some layer-3 products are dead), one node mapped twice."""
    rng = np.random.default_rng(7)
    spec = sus2sh.synthetic_model(2, 2, 6, species_count=2, rb_size=4, max_dist=4.0, seed=17)
    g = spec.graph
    coeffs = g.coeffs * rng.uniform(0.5, 1.5, size=g.coeffs.shape)
    targets = np.unique(g.products[:, 2])
    scalars = rng.choice(np.concatenate([targets, [0, 1]]), size=g.scalar_count).astype(np.int32)
    scalars[1] = scalars[0]
    g2 = dataclasses.replace(g, coeffs=coeffs, scalars=scalars)
    spec2 = dataclasses.replace(spec, graph=g2,
                                moment_coeffs=rng.normal(0.0, 0.3, size=g.scalar_count))
    path = sus2sh.write_mtp(spec2, str(model_dir / "lane_nonstandard_b6.mtp"))
    info = sus2sh.model_check(path)
    assert info["standard_graph"] == 0 and info["layer_count"] == 3
    _check(path, max_slots=1024, n=3)


def test_budget_too_small(make_model):
    path = make_model("lane_budget", l_max=2, k_max=2, body_order=5, rb_size=4, max_dist=4.0,
                      seed=6)
    B = Oracle(path).dims["B"]
    _check(path, max_slots=2 * B + 3, expect_ok=False)
    _check(path, max_slots=2 * B + 40, expect_ok=True)


def test_non_standard_depth2_graph(model_dir):
    """A depth-2 graph that is not the standard one (perturbed coefficients, scalars mapped
    to other nodes, some mapped twice, a mapped basic): the lane program is built from the
    file's graph, not from the standard generator."""
    rng = np.random.default_rng(42)
    spec = sus2sh.synthetic_model(2, 2, 4, species_count=2, rb_size=4, max_dist=4.0, seed=9)
    g = spec.graph
    coeffs = g.coeffs * rng.uniform(0.5, 1.5, size=g.coeffs.shape)
    targets = np.unique(g.products[:, 2])
    scalars = rng.choice(np.concatenate([targets, [0, 1]]), size=g.scalar_count).astype(np.int32)
    g2 = dataclasses.replace(g, coeffs=coeffs, scalars=scalars)
    spec2 = dataclasses.replace(spec, graph=g2,
                                moment_coeffs=rng.normal(0.0, 0.3, size=g.scalar_count))
    path = sus2sh.write_mtp(spec2, str(model_dir / "lane_nonstandard.mtp"))
    assert sus2sh.model_check(path)["standard_graph"] == 0
    _check(path, max_slots=256)
