"""Explicit (non-standard) product graphs -- CPU only: a model whose saved graph computes a
subset of the standard scalars (as a pruned training graph does) loads through the
explicit-graph path, and the oracle's energy equals the standard model's with the dropped
scalars' weights set to zero (same moments, fewer terms), atom by atom."""
import copy
import os

import numpy as np

import sus2sh
from oracle.binding import Oracle


def pruned_pair(model_dir, name, **kw):
    """(standard model with zero weights on the dropped scalars, explicit pruned model)."""
    spec = sus2sh.synthetic_model(**kw)
    g = spec.graph
    # keep body orders <= 3 and every other higher-order scalar
    keep = (g.scalar_body <= 3) | (np.arange(g.scalar_count) % 2 == 0)
    full = copy.deepcopy(spec)
    full.moment_coeffs = np.where(keep, spec.moment_coeffs, 0.0)
    pruned = copy.deepcopy(spec)
    pruned.graph = sus2sh.subgraph(g, keep)
    pruned.moment_coeffs = spec.moment_coeffs[keep]
    a = sus2sh.write_mtp(full, os.path.join(str(model_dir), name + "_full.mtp"))
    b = sus2sh.write_mtp(pruned, os.path.join(str(model_dir), name + "_pruned.mtp"))
    return a, b, g, pruned.graph


MODELS = {
    "l3k2_b5": dict(l_max=3, k_max=2, body_order=5, rb_size=6, max_dist=4.6, seed=91,
                    scaling=0.5),
    "l2k2_b6": dict(l_max=2, k_max=2, body_order=6, rb_size=5, max_dist=4.4, seed=92,
                    scaling=0.5),
}


def test_pruned_graph_loads_as_explicit_and_matches_standard(model_dir):
    from sus2sh import structures as st
    for name, kw in MODELS.items():
        full, pruned, g, h = pruned_pair(model_dir, "expl_" + name, **kw)
        assert h.product_count < g.product_count and h.node_count < g.node_count
        info = sus2sh.model_check(pruned)
        assert info["standard_graph"] == 0
        assert sus2sh.model_check(full)["standard_graph"] == 1
        pos, types, cell = st.fcc_alloy(2, 2, 2, 3.7, [0.5, 0.5], 0.08, 5, 6)
        e0, f0, v0 = Oracle(full).compute(cell, types, pos)
        e1, f1, v1 = Oracle(pruned).compute(cell, types, pos)
        assert np.abs(e0 - e1).max() < 1e-12 * max(1.0, np.abs(e0).max())
        assert np.abs(f0 - f1).max() < 1e-11 * max(1.0, np.abs(f0).max())
        assert np.abs(v0 - v1).max() < 1e-11 * max(1.0, np.abs(v0).max())
