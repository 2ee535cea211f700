"""torch.autograd binding: energy -> (positions, model parameters), FP64 and FP32, checked
against central finite differences (h = 1e-4) and the FP64 CPU oracle's gradients."""
import numpy as np
import pytest

import sus2sh
from sus2sh import structures as st
from oracle.binding import Oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fp32", [False, True])
def test_autograd_positions_and_parameters(make_model, fp32):
    import torch
    from sus2sh.autograd import EnergyModel
    path = make_model("l2k2_autograd", l_max=2, k_max=2, body_order=5, rb_size=6, max_dist=4.6,
                      seed=61, scal_mode="random", scaling=0.5)
    pos, types, cell = st.fcc_alloy(2, 2, 2, 3.6, [0.5, 0.5], 0.1, 7, 8)
    N = len(types)
    print(f"[case] autograd N={N} fp32={fp32}", flush=True)
    e0, f0, _, g0 = Oracle(path).compute(cell, types, pos, param_grad=True)
    model = EnergyModel(path, cell, N, ["Cu", "Zr"], ["sus2_float=%d" % fp32])
    t = torch.as_tensor(types, device="cuda")
    x = torch.tensor(pos, device="cuda", requires_grad=True)
    E = model(x, t)
    E.backward()
    dx = x.grad.cpu().numpy()
    dth = model.params.grad.cpu().numpy()
    tol = 1e-4 if fp32 else 1e-9
    assert abs(E.item() - e0.sum()) <= tol * max(1.0, abs(e0.sum())) * (10 if fp32 else 1)
    assert np.abs(dx + f0).max() <= tol * max(1.0, np.abs(f0).max())
    assert np.abs(dth - g0).max() <= tol * max(1.0, np.abs(g0).max())
    if fp32:
        return
    # central differences (h = 1e-4) of the engine's own energy
    h = 1e-4
    rng = np.random.default_rng(1)
    for _ in range(4):
        i, c = int(rng.integers(N)), int(rng.integers(3))
        xp = x.detach().clone()
        xm = x.detach().clone()
        xp[c, i] += h
        xm[c, i] -= h
        fd = (model(xp, t).item() - model(xm, t).item()) / (2 * h)
        assert fd == pytest.approx(dx[c, i], abs=1e-6 * max(1.0, abs(fd)))
    theta = model.params.detach().clone()
    for k in rng.choice(len(theta), 8, replace=False):
        tp, tm = theta.clone(), theta.clone()
        tp[k] += h
        tm[k] -= h
        with torch.no_grad():
            model.params.copy_(tp)
            ep = model(x.detach(), t).item()
            model.params.copy_(tm)
            em = model(x.detach(), t).item()
            model.params.copy_(theta)
        fd = (ep - em) / (2 * h)
        assert fd == pytest.approx(dth[k], abs=1e-6 * max(1.0, abs(fd))), k
