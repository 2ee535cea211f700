"""GPU: models with explicit (pruned, non-standard) product graphs against the FP64 oracle,
on the default (autotuned) graph kernel and on the lane program, in both precisions."""
import pytest

import sus2sh
from sus2sh import structures as st
from oracle.binding import Oracle

from test_explicit_graph import MODELS, pruned_pair
from test_gpu_parity import _compare

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant", ["auto", "lane"])
@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("name", sorted(MODELS))
def test_pruned_graph_matches_oracle(model_dir, name, fp32, variant):
    _, path, _, h = pruned_pair(model_dir, "gexpl_" + name, **MODELS[name])
    pos, types, cell = st.fcc_alloy(4, 4, 4, 3.7, [0.5, 0.5], 0.08, 15, 16)
    opts = ["Cu", "Zr", "sus2_float=%d" % fp32]
    if variant == "lane":
        opts.append("sus2_sh_graph_variant=3")
    print(f"[case] explicit graph {name} fp32={fp32} {variant} products={h.product_count}",
          flush=True)
    # the pruned body-6 graph needs 384 per-lane slots: 96 KiB per 64-atom tile in FP32,
    # more than the 160 KiB of a workgroup in FP64
    if variant == "lane" and not sus2sh.lane_check(path, max_slots=640 if fp32 else 320,
                                                   waves=4)[0]["ok"]:
        with pytest.raises(sus2sh.Sus2shError) as err:
            sus2sh.evaluate(path, cell, types, pos, options=opts)
        assert "no lane program" in err.value.message
        return
    ref = Oracle(path).compute(cell, types, pos)
    got = sus2sh.evaluate(path, cell, types, pos, options=opts)
    _compare(ref, got, fp32)
