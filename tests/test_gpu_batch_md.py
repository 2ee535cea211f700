"""GPU: batched uneven structures and the NVE driver, against the FP64 oracle."""
import numpy as np
import pytest

import sus2sh
from sus2sh import structures as st
from oracle.binding import Oracle

from test_gpu_parity import _separated_positions

pytestmark = pytest.mark.gpu

MASSES = np.array([63.546, 91.224])  # Cu, Zr (amu)


def _batch(rng, count, nmin, nmax, density=0.065):
    cells, pos, types, offs = [], [], [], [0]
    for _ in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        cell = st.triclinic_cell(n, density, rng)
        cells.append(cell)
        pos.append(_separated_positions(cell, n, 1.8, rng))
        types.append(rng.integers(0, 2, size=n).astype(np.int32))
        offs.append(offs[-1] + n)
    return cells, np.concatenate(pos, axis=1), np.concatenate(types), np.array(offs)


@pytest.mark.parametrize("fp32", [False, True])
def test_batch_matches_per_structure_oracle(make_model, fp32):
    import torch
    path = make_model("l3k3_batch", l_max=3, k_max=3, body_order=5, rb_size=8, max_dist=5.0,
                      seed=51, scaling=0.5)
    rng = np.random.default_rng(11)
    cells, pos, types, offs = _batch(rng, 12, 8, 48)
    N = len(types)
    print(f"[case] batch of {len(cells)} structures, N={N}, fp32={fp32}", flush=True)
    orc = Oracle(path)
    pot = sus2sh.SUS2_SH(path, N, ["Cu", "Zr", "sus2_float=%d" % fp32])
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
    tol = (2e-5, 2e-4, 2e-4) if fp32 else (1e-9, 1e-8, 1e-8)
    for k, cell in enumerate(cells):
        a, b = offs[k], offs[k + 1]
        e0, f0, v0, g0 = orc.compute(cell, types[a:b], pos[:, a:b], brute=True, param_grad=True)
        gref += g0
        assert np.abs(e[a:b] - e0).max() <= tol[0] * max(1, np.abs(e0).max())
        assert np.abs(f[:, a:b] - f0).max() <= tol[1] * max(1, np.abs(f0).max())
        assert np.abs(v[:, a:b] - v0).max() <= tol[2] * max(1, np.abs(v0).max())
    err = np.abs(g - gref).max() / max(1.0, np.abs(gref).max())
    print(f"  param-grad rel err {err:.2e}", flush=True)
    assert err <= (1e-4 if fp32 else 1e-9)
    pot.close()


def _oracle_verlet(orc, cell, types, pos, vel, dt, nsteps):
    m = MASSES[types]
    acc = 9.648533212331e-3
    e, f, _ = orc.compute(cell, types, pos)
    for _ in range(nsteps):
        vel = vel + 0.5 * dt * acc * f / m[None, :]
        pos = pos + dt * vel
        e, f, _ = orc.compute(cell, types, pos)
        vel = vel + 0.5 * dt * acc * f / m[None, :]
    return pos, vel, e


def test_nve_trajectory_and_energy_conservation(make_model):
    import torch
    path = make_model("l2k2_md", l_max=2, k_max=2, body_order=5, rb_size=8, max_dist=5.0,
                      seed=1003, scaling=0.05, pair_core=5.0)
    pos, types, cell = st.fcc_alloy(3, 3, 3, 3.80, [0.5, 0.5], 0.05, 1001, 1002)
    vel = sus2sh.maxwell_velocities(MASSES[types], 300.0, 1004)
    N = len(types)
    print(f"[case] NVE N={N}", flush=True)
    pot = sus2sh.SUS2_SH(path, N, ["Cu", "Zr", "sus2_float=0"])
    t = torch.as_tensor(types, device="cuda")
    x = torch.as_tensor(pos.copy(), device="cuda")
    vv = torch.as_tensor(vel.copy(), device="cuda")
    box = sus2sh.box_struct(cell)
    trace, *_ = sus2sh.md_nve(pot, box, t, x, vv, MASSES, 1.0, 20)
    torch.cuda.synchronize()
    p_ref, v_ref, _ = _oracle_verlet(Oracle(path), cell, types, pos, vel, 1.0, 20)
    assert np.abs(x.cpu().numpy() - p_ref).max() < 1e-9
    assert np.abs(vv.cpu().numpy() - v_ref).max() < 1e-10
    # 100 more steps: total energy drift stays small
    trace, *_ = sus2sh.md_nve(pot, box, t, x, vv, MASSES, 1.0, 100)
    tot = trace.sum(axis=1)
    drift = np.abs(tot - tot[0]).max() / N
    print(f"  max |E_tot - E_tot(0)| / N over 100 fs = {drift:.2e} eV/atom", flush=True)
    assert drift < 1e-4
    pot.close()
