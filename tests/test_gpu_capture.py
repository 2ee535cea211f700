"""GPU: host-synchronisation-free evaluations (sus2_sh_deferred_check=1), their capture in a
HIP graph, the deferred capacity check, and graph-replayed NVE steps, each against the
synchronised eager path of the same engine."""
import numpy as np
import pytest

import sus2sh
from sus2sh import structures as st

pytestmark = pytest.mark.gpu

MASSES = np.array([63.546, 91.224])  # Cu, Zr (amu)


def _model(make_model):
    return make_model("l3k3_capture", l_max=3, k_max=3, body_order=5, rb_size=8, max_dist=5.0,
                      seed=71, scaling=0.5)


def _arrays(types, pos):
    import torch
    n = len(types)
    t = torch.as_tensor(types, device="cuda")
    x = torch.as_tensor(np.ascontiguousarray(pos), device="cuda")
    e = torch.zeros(n, dtype=torch.float64, device="cuda")
    f = torch.zeros(3, n, dtype=torch.float64, device="cuda")
    v = torch.zeros(9, n, dtype=torch.float64, device="cuda")
    return t, x, e, f, v


def _eager(path, cell, types, pos, fp32):
    import torch
    pot = sus2sh.SUS2_SH(path, len(types), ["Cu", "Zr", "sus2_float=%d" % fp32])
    t, x, e, f, v = _arrays(types, pos)
    pot.compute(sus2sh.box_struct(cell), t, x, e, f, v)
    torch.cuda.synchronize()
    pot.close()
    return e.cpu().numpy(), f.cpu().numpy(), v.cpu().numpy()


def _same(got, ref, fp32):
    """Same kernels on the same inputs; adjoints and neighbour forces are summed with
    atomics in arbitrary order, so results agree to rounding: 1e-5 (FP32 accumulators) and
    1e-11 (FP64) in eV, eV/A and eV."""
    tol = 1e-5 if fp32 else 1e-11
    for a, b in zip(got, ref):
        assert np.abs(a - b).max() <= tol, float(np.abs(a - b).max())


@pytest.mark.parametrize("fp32", [True, False])
def test_deferred_calls_match_synchronised(make_model, fp32):
    import torch
    path = _model(make_model)
    pos, types, cell = st.fcc_alloy(6, 6, 6, 3.85, [0.5, 0.5], 0.08, 7, 8)
    print(f"[case] deferred check, N={len(types)}, fp32={fp32}", flush=True)
    ref = _eager(path, cell, types, pos, fp32)
    pot = sus2sh.SUS2_SH(path, len(types), ["Cu", "Zr", "sus2_float=%d" % fp32,
                                            "sus2_sh_deferred_check=1"])
    t, x, e, f, v = _arrays(types, pos)
    box = sus2sh.box_struct(cell)
    for _ in range(3):  # synchronised sizing call, then two deferred calls
        e.zero_(), f.zero_(), v.zero_()
        pot.compute(box, t, x, e, f, v)
        torch.cuda.synchronize()
        _same((e.cpu().numpy(), f.cpu().numpy(), v.cpu().numpy()), ref, fp32)
    pot.check()
    assert abs(pot.mean_neighbors() - 42) < 3
    pot.close()


def test_deferred_capacity_overflow_is_reported(make_model):
    """A deferred call on a denser structure (same atoms and box, positions pulled towards
    the origin) exceeds the capacity sized by the first call: the kernels clip the lists,
    check() reports status 5, and the next call (synchronised again) is exact."""
    import torch
    path = _model(make_model)
    pos, types, cell = st.fcc_alloy(6, 6, 6, 3.85, [0.5, 0.5], 0.08, 7, 8)
    dense = 0.8 * pos  # half-empty box, 1.95x the density in the occupied part
    print(f"[case] capacity overflow, N={len(types)}", flush=True)
    ref = _eager(path, cell, types, dense, True)
    pot = sus2sh.SUS2_SH(path, len(types), ["Cu", "Zr", "sus2_sh_deferred_check=1"])
    t, x, e, f, v = _arrays(types, pos)
    box = sus2sh.box_struct(cell)
    pot.compute(box, t, x, e, f, v)
    cap0 = pot.info()["neighbor_capacity"]
    x.copy_(torch.as_tensor(np.ascontiguousarray(dense), device="cuda"))
    e.zero_(), f.zero_(), v.zero_()
    pot.compute(box, t, x, e, f, v)  # deferred, overflows
    with pytest.raises(sus2sh.Sus2shError) as err:
        pot.check()
    assert err.value.code == 5
    assert pot.info()["neighbor_capacity"] > cap0
    pot.check()  # the latch was cleared
    e.zero_(), f.zero_(), v.zero_()
    pot.compute(box, t, x, e, f, v)  # synchronised sizing call
    torch.cuda.synchronize()
    _same((e.cpu().numpy(), f.cpu().numpy(), v.cpu().numpy()), ref, True)
    pot.close()


@pytest.mark.parametrize("fp32", [True, False])
def test_graph_capture_replays_match_eager(make_model, fp32):
    """[zero outputs, compute] captured once and replayed with eager work and
    synchronisations in between; every replay equals the eager result, also after the
    positions buffer is overwritten in place."""
    import torch
    path = _model(make_model)
    pos, types, cell = st.fcc_alloy(6, 6, 6, 3.85, [0.5, 0.5], 0.08, 7, 8)
    rng = np.random.default_rng(5)
    pos2 = pos + rng.normal(0.0, 0.03, size=pos.shape)
    print(f"[case] graph capture, N={len(types)}, fp32={fp32}", flush=True)
    refs = [_eager(path, cell, types, p, fp32) for p in (pos, pos2)]
    pot = sus2sh.SUS2_SH(path, len(types), ["Cu", "Zr", "sus2_float=%d" % fp32,
                                            "sus2_sh_deferred_check=1"])
    t, x, e, f, v = _arrays(types, pos)
    box = sus2sh.box_struct(cell)

    def step():
        e.zero_()
        f.zero_()
        v.zero_()
        pot.compute(box, t, x, e, f, v)

    step()
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        step()
    torch.cuda.synchronize()
    for k, p in enumerate((pos, pos, pos2, pos2, pos)):
        x.copy_(torch.as_tensor(np.ascontiguousarray(p), device="cuda"))
        junk = torch.ones(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = (e.cpu().numpy(), f.cpu().numpy(), v.cpu().numpy())
        _same(got, refs[0] if k in (0, 1, 4) else refs[1], fp32)
        assert float(junk) == float(1 << 20)
    pot.check()
    del graph
    pot.close()


@pytest.mark.parametrize("fp32", [True, False])
def test_md_graph_replay_matches_eager_steps(make_model, fp32):
    """sus2sh_md_nve with captured steady steps (two chunks of 50 and a partial one) gives
    the same trajectory and energy trace as the step-by-step eager driver (to the rounding
    of the order of the force atomics)."""
    import torch
    path = make_model("l2k2_md_graph", l_max=2, k_max=2, body_order=5, rb_size=8,
                      max_dist=5.0, seed=1003, scaling=0.05, pair_core=5.0)
    pos, types, cell = st.fcc_alloy(5, 5, 5, 3.80, [0.5, 0.5], 0.05, 1001, 1002)
    vel = sus2sh.maxwell_velocities(MASSES[types], 300.0, 1004)
    print(f"[case] md graph vs eager, N={len(types)}, fp32={fp32}", flush=True)
    out = []
    for graph in (0, 1):
        pot = sus2sh.SUS2_SH(path, len(types), ["Cu", "Zr", "sus2_float=%d" % fp32,
                                                "sus2_sh_md_graph=%d" % graph])
        t = torch.as_tensor(types, device="cuda")
        x = torch.as_tensor(np.ascontiguousarray(pos), device="cuda")
        vv = torch.as_tensor(np.ascontiguousarray(vel), device="cuda")
        trace, *_ = sus2sh.md_nve(pot, sus2sh.box_struct(cell), t, x, vv, MASSES, 1.0, 120)
        torch.cuda.synchronize()
        out.append((trace, x.cpu().numpy(), vv.cpu().numpy()))
        pot.close()
    (tr0, x0, v0), (tr1, x1, v1) = out
    assert np.abs(x1 - x0).max() <= 1e-6
    assert np.abs(v1 - v0).max() <= 1e-8
    assert np.abs(tr1 - tr0).max() <= 1e-6 * np.abs(tr0).max()
    tot = out[1][0].sum(axis=1)
    assert np.abs(tot - tot[0]).max() / len(types) < 1e-3
