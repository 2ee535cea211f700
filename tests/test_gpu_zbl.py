"""GPU: SUS2-SH models with the ZBL pair term (zbl_enabled), against the FP64 oracle with
the same pair tables -- single structures (global and type-wise cutoffs, both precisions),
a batch, and the parameter-gradient path."""
import os

import numpy as np
import pytest

import sus2sh
from oracle.binding import Oracle

from test_gpu_parity import _compare, _separated_positions

pytestmark = pytest.mark.gpu


def _model(model_dir, name, extra):
    spec = sus2sh.synthetic_model(2, 2, 5, species_count=2, rb_size=8, max_dist=5.0, seed=41,
                                  scaling=0.05, pair_core=5.0)
    spec.extra = dict(extra)
    return sus2sh.write_mtp(spec, os.path.join(str(model_dir), name + ".mtp"))


def _close_structure(seed, n=120, density=0.07, dmin=0.95):
    rng = np.random.default_rng(seed)
    L = (n / density) ** (1 / 3)
    cell = np.diag([L, L, L])
    pos = _separated_positions(cell, n, dmin, rng)
    types = rng.integers(0, 2, size=n).astype(np.int32)
    return pos, types, cell


def _oracle(path, symbols):
    z, inner, outer = sus2sh.zbl_tables(path, symbols)
    orc = Oracle(path)
    orc.set_zbl(z, inner, outer)
    return orc, outer


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("typewise", [False, True])
def test_zbl_single_structure_matches_oracle(model_dir, fp32, typewise):
    extra = {"zbl_enabled": True, "zbl_inner": 0.8, "zbl_outer": 1.6}
    if typewise:
        extra["zbl_typewise_cutoff_factor"] = 0.5
    path = _model(model_dir, "zbl_tw%d" % typewise, extra)
    pos, types, cell = _close_structure(5)
    orc, outer = _oracle(path, ["Cu", "Zr"])
    ref = orc.compute(cell, types, pos)
    # the structure must exercise the pair term
    d = pos[:, :, None] - pos[:, None, :]
    d -= np.round(d / cell[0, 0]) * cell[0, 0]
    r = np.sqrt((d * d).sum(0)) + np.eye(len(types)) * 1e9
    assert (r < outer.min()).sum() >= 4
    got = sus2sh.evaluate(path, cell, types, pos, options=["Cu", "Zr", "sus2_float=%d" % fp32])
    print(f"[case] zbl typewise={typewise} fp32={fp32} N={len(types)}", flush=True)
    _compare(ref, got, fp32)


def test_zbl_batch_and_param_grad(model_dir):
    import torch
    path = _model(model_dir, "zbl_batch", {"zbl_enabled": True, "zbl_inner": 0.8,
                                           "zbl_outer": 1.6})
    orc, _ = _oracle(path, ["Cu", "Zr"])
    structs = [_close_structure(s, n=40 + 20 * s) for s in range(3)]
    cells = [c for _, _, c in structs]
    pos = np.concatenate([p for p, _, _ in structs], axis=1)
    types = np.concatenate([t for _, t, _ in structs])
    offs = np.cumsum([0] + [len(t) for _, t, _ in structs])
    N = len(types)
    print(f"[case] zbl batch N={N}", flush=True)
    pot = sus2sh.SUS2_SH(path, N, ["Cu", "Zr", "sus2_float=0"])
    P = pot.info()["param_count"]
    t = torch.as_tensor(types, device="cuda")
    x = torch.as_tensor(np.ascontiguousarray(pos), device="cuda")
    e = torch.zeros(N, dtype=torch.float64, device="cuda")
    f = torch.zeros(3, N, dtype=torch.float64, device="cuda")
    v = torch.zeros(9, N, dtype=torch.float64, device="cuda")
    g = torch.zeros(P, dtype=torch.float64, device="cuda")
    sus2sh.compute_batch(pot, cells, offs, t, x, e, f, v, param_grad=g)
    torch.cuda.synchronize()
    e, f, v, g = e.cpu().numpy(), f.cpu().numpy(), v.cpu().numpy(), g.cpu().numpy()
    gref = np.zeros(P)
    for k, cell in enumerate(cells):
        a, b = offs[k], offs[k + 1]
        e0, f0, v0, g0 = orc.compute(cell, types[a:b], pos[:, a:b], brute=True, param_grad=True)
        gref += g0
        _compare((e0, f0, v0), (e[a:b], f[:, a:b], v[:, a:b]), False)
    assert np.abs(g - gref).max() <= 1e-9 * max(1.0, np.abs(gref).max())
    pot.close()
