"""GPU parity of body-order-6 (three-layer) graphs on the lane kernel, through the C ABI.

The reference's production q-total model is body order 6 with sh_body_l_max = {4,4,4,4,1}
(docs/gpumd-sus2-sh-interface-plan.md:2041-2045): its layer-3 products contract a layer-2
tensor of l <= 1 (a body-5 intermediate) with a layer-1 node.  The lane program keeps those
layer-2 tensors (Y) in its quadratic form and their layer-1 operands (P) in persistent
per-lane slots.  Each case pins graph variant 3 and compares E, F, W with the FP64 oracle's
flat graph evaluation (src/force/sus2_sh.cu:4799-4846 semantics) on a 4,608-atom piece of
the benchmark structure (12x12x8 fcc cells, a = 3.85 A, jitter 0.08 A), then leaves the
autotuner on (it runs from 4,096 atoms) and checks the kernel it keeps.

Tolerances as in test_gpu_parity: FP64 1e-9 eV/atom, 1e-8 eV/A, 1e-8 eV;
FP32 2e-5 eV/atom + 2e-6 |E|, 1e-4 eV/A + 2e-5 |F|max, 1e-4 eV + 2e-5 |W|max.
"""
import numpy as np
import pytest

import sus2sh
from sus2sh import structures as st
from oracle.binding import Oracle

from test_gpu_lane_shapes import _compare, _evaluate

pytestmark = pytest.mark.gpu

MODELS = {
    # production shape at l_max 3 / 4 (73 Y nodes, 45 P nodes in both)
    "l3k3_33331": dict(l_max=3, k_max=3, body_l_max=(3, 3, 3, 3, 1), seed=601, scaling=0.4),
    "l4k3_44441": dict(l_max=4, k_max=3, body_l_max=(4, 4, 4, 4, 1), seed=602, scaling=0.3),
}


@pytest.fixture(scope="module")
def body6_cases(model_dir):
    pos, types, cell = st.fcc_alloy(12, 12, 8, 3.85, [0.5, 0.5], 0.08, 3001, 3002)
    out = {}
    for name, kw in MODELS.items():
        spec = sus2sh.synthetic_model(body_order=6, species_count=2, rb_size=8, max_dist=5.0,
                                      **kw)
        path = sus2sh.write_mtp(spec, str(model_dir / f"body6_{name}.mtp"))
        out[name] = (path, Oracle(path).compute(cell, types, pos))
    return pos, types, cell, out


SHAPES = [(4, 4), (8, 2), (4, 2), (2, 4), (16, 2)]


@pytest.mark.parametrize("fp32", [True, False])
@pytest.mark.parametrize("waves,nslot", SHAPES)
@pytest.mark.parametrize("name", sorted(MODELS))
def test_body6_lane_kernel_matches_oracle(body6_cases, name, waves, nslot, fp32):
    pos, types, cell, cases = body6_cases
    path, ref = cases[name]
    assert sus2sh.model_check(path)["layer_count"] == 3
    # the smallest slot budget the program fits; FP64 tiles are limited to 320 slots
    # (160 KiB), which the l4k3 program (384 slots) exceeds
    fits = sus2sh.lane_check(path, max_slots=640 if fp32 else 320, waves=waves,
                             nslot=nslot)[0]["ok"]
    opts = ["Cu", "Zr", f"sus2_float={int(fp32)}", "sus2_sh_graph_variant=3",
            f"sus2_sh_lane_shape={waves}x{nslot}x160"]
    print(f"[case] body6 {name} lane ({waves},{nslot}) fp32={fp32} N={len(types)} fits={fits}",
          flush=True)
    if not fits:
        with pytest.raises(sus2sh.Sus2shError) as err:
            _evaluate(path, cell, types, pos, opts)
        assert "no lane program" in err.value.message
        return
    got, info = _evaluate(path, cell, types, pos, opts)
    assert info["graph_variant"] == 3
    assert (info["lane_waves"], info["lane_nslot"]) == (waves, nslot)
    print(f"  lane slots {info['lane_slots']}", flush=True)
    _compare(ref, got, fp32)


@pytest.mark.parametrize("fp32", [True, False])
@pytest.mark.parametrize("name", sorted(MODELS))
def test_body6_autotuned_matches_oracle(body6_cases, name, fp32):
    pos, types, cell, cases = body6_cases
    path, ref = cases[name]
    print(f"[case] body6 {name} autotuned fp32={fp32} N={len(types)}", flush=True)
    got, info = _evaluate(path, cell, types, pos, ["Cu", "Zr", f"sus2_float={int(fp32)}"])
    print(f"  kept variant {info['graph_variant']} tile {info['graph_tile']} lane shape "
          f"({info['lane_waves']},{info['lane_nslot']},{info['lane_slots']},"
          f"{info['lane_atoms']})", flush=True)
    _compare(ref, got, fp32)
