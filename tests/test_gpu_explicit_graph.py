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
    if variant == "lane" and MODELS[name]["body_order"] > 5:
        pytest.skip("the lane program covers graphs of at most two product layers")
    pos, types, cell = st.fcc_alloy(4, 4, 4, 3.7, [0.5, 0.5], 0.08, 15, 16)
    opts = ["Cu", "Zr", "sus2_float=%d" % fp32]
    if variant == "lane":
        opts.append("sus2_sh_graph_variant=3")
    print(f"[case] explicit graph {name} fp32={fp32} {variant} products={h.product_count}",
          flush=True)
    ref = Oracle(path).compute(cell, types, pos)
    got = sus2sh.evaluate(path, cell, types, pos, options=opts)
    _compare(ref, got, fp32)
