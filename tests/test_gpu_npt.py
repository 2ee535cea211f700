"""NPT (MTTK barostat + Nose-Hoover chains) around the potential on the GPU (through the C ABI).

The reference's benchmark protocol runs GPUMD's `ensemble npt_mttk temp T1 T2 aniso p1 p2
tperiod 50 pperiod 500` (docs/gpumd-sus2-sh-interface-plan.md:264-271); the integrator itself
is not in the reference overlay, so sus2sh_md_npt_mttk is checked through what the published
scheme fixes:

* its conserved extended energy (potential + kinetic + thermostat and barostat chains +
  barostat kinetic energy + P0 (V - V0)) stays flat over 800 steps at dt 0.5 fs, with the
  target pressure 1 GPa below the start so the box has to move (a wrong sign or unit in the
  barostat force, or a virial of the wrong sign, breaks it);
* with thermostat and barostat periods of 1e12 steps it is velocity Verlet: positions and
  velocities equal sus2sh_md_nve's after 40 steps, and the box does not move;
* the step-0 pressure and temperature it reports equal (sum m v^2 + sum W_aa) / 3V and
  sum m v^2 / (3N - 3) k_B from the potential's own virial and the velocities, evaluated
  here on the host (units: GPa, K);
* iso keeps the box's axis ratios; the box inverse stays the inverse.

Model and structure: the C2 MD model (Cu-Zr l3k3 with a repulsive pair core, rb 8, rc 5.0)
on a 6x6x6-cell fcc piece (864 atoms, 64/36, jitter 0.08), 300 K Maxwell velocities.
"""
import numpy as np
import pytest

import sus2sh
from sus2sh import configs

pytestmark = pytest.mark.gpu

KB = 8.617333262e-5            # eV/K
EV_PER_A3_PER_GPA = 1.0 / 160.21766208
MV2_TO_EV = 1.0 / 9.648533212331e-3   # amu A^2/fs^2 -> eV


def _dev(a, dtype):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=dtype), device="cuda")


@pytest.fixture(scope="module")
def npt_case(model_dir):
    path = sus2sh.write_mtp(configs.model_c2(md=True), str(model_dir / "c2_md_npt.mtp"))
    pos, types, cell = configs.structure_c2(6)
    masses = np.array([configs.MASSES["Cu"], configs.MASSES["Zr"]])
    vel = sus2sh.maxwell_velocities(masses[types], 300.0, 1006)
    return path, pos, types, cell, masses, vel


def _run_npt(case, fp32, nsteps, dt, p_shift, tperiod=50.0, pperiod=500.0, couple="aniso",
             p_target=None):
    import torch
    path, pos, types, cell, masses, vel = case
    n = len(types)
    pot = sus2sh.SUS2_SH(path, n, ["Cu", "Zr", f"sus2_float={int(fp32)}"])
    box = sus2sh.box_struct(cell)
    t, x, v = _dev(types, np.int32), _dev(pos, np.float64), _dev(vel, np.float64)
    if p_target is None:
        tr0, *_ = sus2sh.md_npt_mttk(pot, box, t, x.clone(), v.clone(), masses, dt, 0, 300.0,
                                     300.0, 0.0, 0.0, tperiod, pperiod, couple)
        p_target = tr0[0, 5] + p_shift
    trace, e, f, w = sus2sh.md_npt_mttk(pot, box, t, x, v, masses, dt, nsteps, 300.0, 300.0,
                                        p_target, p_target, tperiod, pperiod, couple)
    torch.cuda.synchronize()
    pot.close()
    return trace, box, x.cpu().numpy(), v.cpu().numpy(), w.cpu().numpy(), p_target


def _box_matrix(box):
    h = np.array(list(box.h))
    return h[:9].reshape(3, 3), h[9:].reshape(3, 3)


def test_step0_pressure_and_temperature(npt_case):
    """Reported P and T at step 0 equal the host's evaluation from the virial and velocities."""
    path, pos, types, cell, masses, vel = npt_case
    n = len(types)
    print(f"[case] NPT step 0 N={n}", flush=True)
    trace, box, _, _, w, _ = _run_npt(npt_case, False, 0, 0.5, 0.0, p_target=0.0)
    m = masses[types]
    mv2 = (m[None, :] * vel ** 2).sum(axis=1) * MV2_TO_EV   # per axis, eV (vel is [3, N])
    vol = abs(np.linalg.det(cell))
    p_host = (mv2.sum() + w[0:3].sum()) / (3.0 * vol) / EV_PER_A3_PER_GPA
    t_host = mv2.sum() / ((3 * n - 3) * KB)
    print(f"  P {trace[0, 5]:.6f} GPa (host {p_host:.6f}), T {trace[0, 4]:.3f} K "
          f"(host {t_host:.3f}), V {trace[0, 3]:.3f} (host {vol:.3f})", flush=True)
    assert abs(trace[0, 5] - p_host) <= 1e-9 * max(1.0, abs(p_host))
    assert abs(trace[0, 4] - t_host) <= 1e-9 * t_host
    assert abs(trace[0, 3] - vol) <= 1e-9 * vol
    assert abs(trace[0, 6] - np.cbrt(vol)) <= 1e-9 * np.cbrt(vol)
    assert trace[0, 7] == pytest.approx(300.0, rel=1e-12)
    assert abs(trace[0, 1] - 0.5 * mv2.sum()) <= 1e-9 * mv2.sum()


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("couple", ["aniso", "iso"])
def test_conserved_extended_energy(npt_case, fp32, couple):
    n = len(npt_case[2])
    print(f"[case] NPT {couple} fp32={fp32} N={n} 800 steps dt 0.5", flush=True)
    trace, box, *_ = _run_npt(npt_case, fp32, 800, 0.5, -1.0, couple=couple)
    h = trace[:, 2]
    drift = np.abs(h - h[0]).max() / n
    vol = trace[:, 3]
    dv = np.abs(vol - vol[0]).max() / vol[0]
    t_mean = trace[400:, 4].mean()
    print(f"  max|H - H0|/N = {drift:.2e} eV/atom, max|V - V0|/V0 = {dv:.3e}, "
          f"P {trace[0, 5]:.3f} -> {trace[-1, 5]:.3f} GPa (target {trace[0, 5] - 1.0:.3f}), "
          f"mean T (second half) {t_mean:.1f} K", flush=True)
    assert np.all(np.isfinite(trace))
    assert drift < 1e-4
    assert dv > 1e-3          # the barostat moved the box
    assert 150.0 < t_mean < 450.0
    H, Hi = _box_matrix(box)
    assert np.abs(H @ Hi - np.eye(3)).max() < 1e-12
    assert np.abs(H - np.diag(np.diag(H))).max() == 0.0
    if couple == "iso":
        d = np.diag(H)
        c = np.diag(npt_case[3])
        assert np.abs(d / d[0] - c / c[0]).max() < 1e-12


def test_long_periods_reduce_to_velocity_verlet(npt_case):
    import torch
    path, pos, types, cell, masses, vel = npt_case
    n = len(types)
    print(f"[case] NPT periods 1e12 vs NVE N={n} 40 steps", flush=True)
    trace, box, x1, v1, _, _ = _run_npt(npt_case, False, 40, 0.5, 0.0, tperiod=1e12,
                                        pperiod=1e12, p_target=0.0)
    pot = sus2sh.SUS2_SH(path, n, ["Cu", "Zr", "sus2_float=0"])
    t, x, v = _dev(types, np.int32), _dev(pos, np.float64), _dev(vel, np.float64)
    tr_nve, *_ = sus2sh.md_nve(pot, sus2sh.box_struct(cell), t, x, v, masses, 0.5, 40)
    torch.cuda.synchronize()
    pot.close()
    x2, v2 = x.cpu().numpy(), v.cpu().numpy()
    dx, dvel = np.abs(x1 - x2).max(), np.abs(v1 - v2).max()
    H, _ = _box_matrix(box)
    print(f"  max|dx| {dx:.2e} A, max|dv| {dvel:.2e} A/fs, box change "
          f"{np.abs(H - cell.T).max():.2e}", flush=True)
    assert dx < 1e-9 and dvel < 1e-10
    assert np.abs(H - cell.T).max() == 0.0
    assert np.abs(trace[:, 0] - tr_nve[:, 0]).max() < 1e-7
    assert np.abs(trace[:, 1] - tr_nve[:, 1]).max() < 1e-7


def test_rejects_bad_arguments(npt_case):
    path, pos, types, cell, masses, vel = npt_case
    n = len(types)
    pot = sus2sh.SUS2_SH(path, n, ["Cu", "Zr"])
    t, x, v = _dev(types, np.int32), _dev(pos, np.float64), _dev(vel, np.float64)
    with pytest.raises(sus2sh.Sus2shError, match="periodic"):
        sus2sh.md_npt_mttk(pot, sus2sh.box_struct(cell, (1, 1, 0)), t, x, v, masses, 0.5, 1,
                           300.0, 300.0, 0.0, 0.0)
    with pytest.raises(sus2sh.Sus2shError, match="positive"):
        sus2sh.md_npt_mttk(pot, sus2sh.box_struct(cell), t, x, v, masses, 0.5, 1, 300.0,
                           300.0, 0.0, 0.0, tperiod=0.0)
    with pytest.raises(ValueError):
        sus2sh.md_npt_mttk(pot, sus2sh.box_struct(cell), t, x, v, masses, 0.5, 1, 300.0,
                           300.0, 0.0, 0.0, couple="xyz")
    pot.close()
