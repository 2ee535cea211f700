"""GPU: models with the RBLaguerre_log1p radial basis against the FP64 oracle (both
precisions, per-(mu, pair) maps and a uniform map), and the clean refusal of dE/dtheta for
them."""
import numpy as np
import pytest

import sus2sh
from sus2sh import structures as st
from oracle.binding import Oracle

from test_gpu_parity import _compare

pytestmark = pytest.mark.gpu

CASES = {
    "l2k2_random": dict(l_max=2, k_max=2, body_order=5, rb_size=7, max_dist=5.0, seed=61,
                        scal_mode="random", scaling=0.5),
    "l3k3_uniform": dict(l_max=3, k_max=3, body_order=5, rb_size=8, max_dist=5.0, seed=62,
                         scaling=0.5),
}


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("name", sorted(CASES))
def test_laguerre_matches_oracle(make_model, name, fp32):
    path = make_model("lag_" + name, radial_basis_type="RBLaguerre_log1p", **CASES[name])
    pos, types, cell = st.fcc_alloy(4, 4, 4, 3.85, [0.5, 0.5], 0.08, 21, 22)
    print(f"[case] laguerre {name} fp32={fp32} N={len(types)}", flush=True)
    ref = Oracle(path).compute(cell, types, pos)
    got = sus2sh.evaluate(path, cell, types, pos, options=["Cu", "Zr", "sus2_float=%d" % fp32])
    _compare(ref, got, fp32)


def test_laguerre_param_grad_is_refused(make_model):
    import torch
    path = make_model("lag_l2k2_random", radial_basis_type="RBLaguerre_log1p",
                      **CASES["l2k2_random"])
    pos, types, cell = st.fcc_alloy(3, 3, 3, 3.85, [0.5, 0.5], 0.08, 21, 22)
    n = len(types)
    pot = sus2sh.SUS2_SH(path, n, ["Cu", "Zr"])
    t = torch.as_tensor(types, device="cuda")
    x = torch.as_tensor(np.ascontiguousarray(pos), device="cuda")
    e = torch.zeros(n, dtype=torch.float64, device="cuda")
    f = torch.zeros(3, n, dtype=torch.float64, device="cuda")
    v = torch.zeros(9, n, dtype=torch.float64, device="cuda")
    g = torch.zeros(pot.info()["param_count"], dtype=torch.float64, device="cuda")
    with pytest.raises(sus2sh.Sus2shError) as err:
        pot.compute(sus2sh.box_struct(cell), t, x, e, f, v, param_grad=g)
    assert err.value.code == 4
    pot.compute(sus2sh.box_struct(cell), t, x, e, f, v)  # the handle stays usable
    torch.cuda.synchronize()
    pot.close()


def test_reference_smoke_inputs(model_dir):
    """The reference's GPUMD smoke inputs (2-atom H-C cell, eight-symbol potential line with
    sus2_float=1) through the run.in / XYZ readers and the engine, against the oracle; the
    model file the line names is replaced by a synthetic eight-species Laguerre model."""
    import os
    from sus2sh import gpumd_io
    gold = os.path.join(os.path.dirname(__file__), "golden", "gpumd_smoke")
    spec = sus2sh.synthetic_model(2, 2, 5, species_count=8, rb_size=6, max_dist=5.0, seed=81,
                                  radial_basis_type="RBLaguerre_log1p", scaling=0.5)
    mtp = sus2sh.write_mtp(spec, os.path.join(str(model_dir), "smoke_lag8_gpu.mtp"))
    path, opts, cell, pbc, types, pos = gpumd_io.gpumd_inputs(
        os.path.join(gold, "run.in"), os.path.join(gold, "model.xyz"), mtp)
    print(f"[case] gpumd smoke inputs options={opts}", flush=True)
    ref = Oracle(path).compute(cell, types, pos, pbc=pbc)
    got = sus2sh.evaluate(path, cell, types, pos, options=opts, pbc=pbc)
    assert np.abs(ref[1]).max() > 1e-6  # the H-C pair interacts
    _compare(ref, got, True)
