"""GPU checks of the benchmark configurations at bounded sizes (through the C ABI).

C1  Cu-Zr l2k2 (rb 8, rc 5.0) on the full 10x10x10-cell fcc (4,000 atoms): E/F/W of both
    precisions against the FP64 oracle.
C2  Cu-Zr l3k3 (rb 8, rc 5.0, seed 2003) on an 8x8x8-cell piece (2,048 atoms, 64/36, jitter
    0.08): 400 velocity-Verlet steps at dt 0.5 fs from 300 K in both precisions; the total
    energy stays within 1e-4 eV/atom of its start (energy drift leg of C2).
C4  Cu-Zr-Al l4k4 (rb 10, RBChebyshev_sss_rational, rc 5.5, S = 3, 45/45/10) on 3,000 atoms
    at random positions (density 0.060 A^-3, minimum separation 2.0 A): E/F/W and dE/dtheta
    over every model coefficient in both precisions, against the FP64 oracle's analytic
    gradient, and the FP64 engine's forces and dE/dtheta against central differences of its
    own total energy (h = 1e-4) on 6 coordinates and 8 parameters.  The rational map's
    parameter adjoint (xi = z / sqrt(1 + z^2)) runs here.
C5  48 cells of the batched workload's generator (triclinic, 8..256 atoms, multi-image) in
    one batched call with dE/dtheta, both precisions, against the oracle cell by cell.

Tolerances: FP64 E 1e-9 eV/atom, F 1e-8 eV/A, W 1e-8 eV, dE/dtheta 1e-9 of max |dE/dtheta|;
FP32 E 2e-5 eV/atom + 2e-6 |E|, F 1e-4 eV/A + 2e-5 |F|max, W 1e-4 eV + 2e-5 |W|max,
dE/dtheta 1e-4 of max |dE/dtheta|; central differences 1e-5 eV/A (force) and 1e-5 relative
(dE/dtheta): h^2 truncation plus the rounding of total energies of ~1e3 eV, far below the
O(1) relative error of a wrong term.
"""
import numpy as np
import pytest

import sus2sh
from sus2sh import configs
from oracle.binding import Oracle

pytestmark = pytest.mark.gpu


def _dev(a, dtype):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=dtype), device="cuda")


def _engine(path, n, S, fp32):
    syms = ["Cu", "Zr", "Al"][:S]
    return sus2sh.SUS2_SH(path, n, syms + [f"sus2_float={int(fp32)}"])


def _eval(pot, cell, types, pos, param_grad=False):
    import torch
    n = len(types)
    t, x = _dev(types, np.int32), _dev(pos, np.float64)
    e = torch.zeros(n, dtype=torch.float64, device="cuda")
    f = torch.zeros(3, n, dtype=torch.float64, device="cuda")
    v = torch.zeros(9, n, dtype=torch.float64, device="cuda")
    g = torch.zeros(pot.info()["param_count"], dtype=torch.float64, device="cuda") \
        if param_grad else None
    pot.compute(sus2sh.box_struct(cell), t, x, e, f, v, param_grad=g)
    torch.cuda.synchronize()
    out = (e.cpu().numpy(), f.cpu().numpy(), v.cpu().numpy())
    return out + (g.cpu().numpy(),) if param_grad else out


def _check_efw(ref, got, fp32):
    e0, f0, v0 = ref[:3]
    e1, f1, v1 = got[:3]
    de, df, dv = np.abs(e1 - e0).max(), np.abs(f1 - f0).max(), np.abs(v1 - v0).max()
    print(f"  fp32={fp32}: max|dE|={de:.3e} eV/atom max|dF|={df:.3e} eV/A max|dW|={dv:.3e} "
          f"(|F|max={np.abs(f0).max():.3f})", flush=True)
    if fp32:
        assert de <= 2e-5 + 2e-6 * np.abs(e0).max()
        assert df <= 1e-4 + 2e-5 * np.abs(f0).max()
        assert dv <= 1e-4 + 2e-5 * np.abs(v0).max()
    else:
        assert de <= 1e-9 and df <= 1e-8 and dv <= 1e-8


def test_c1_full_size_matches_oracle(model_dir):
    path = sus2sh.write_mtp(configs.model_c1(), str(model_dir / "c1.mtp"))
    pos, types, cell = configs.structure_c1(10)
    print(f"[case] C1 N={len(types)}", flush=True)
    ref = Oracle(path).compute(cell, types, pos)
    for fp32 in (False, True):
        pot = _engine(path, len(types), 2, fp32)
        _check_efw(ref, _eval(pot, cell, types, pos), fp32)
        pot.close()


@pytest.mark.parametrize("fp32", [False, True])
def test_c2_nve_energy_drift(model_dir, fp32):
    import torch
    path = sus2sh.write_mtp(configs.model_c2(md=True), str(model_dir / "c2_md.mtp"))
    pos, types, cell = configs.structure_c2(8)
    n = len(types)
    masses = np.array([configs.MASSES["Cu"], configs.MASSES["Zr"]])
    vel = sus2sh.maxwell_velocities(masses[types], 300.0, 1004)
    print(f"[case] C2 NVE N={n} fp32={fp32} 400 steps", flush=True)
    pot = _engine(path, n, 2, fp32)
    t, x, vv = _dev(types, np.int32), _dev(pos, np.float64), _dev(vel, np.float64)
    trace, *_ = sus2sh.md_nve(pot, sus2sh.box_struct(cell), t, x, vv, masses, 0.5, 400)
    torch.cuda.synchronize()
    pot.close()
    tot = trace.sum(axis=1)
    dev = np.abs(tot - tot[0]).max() / n
    temp = trace[-1, 1] / (1.5 * n * 8.617333262e-5)
    print(f"  max|E_tot - E_tot(0)|/N = {dev:.2e} eV/atom over 0.2 ps, T_end = {temp:.0f} K",
          flush=True)
    assert np.all(np.isfinite(tot))
    assert dev < 1e-4
    assert 50.0 < temp < 600.0


@pytest.fixture(scope="module")
def c4_case(model_dir):
    path = sus2sh.write_mtp(configs.model_c4(), str(model_dir / "c4.mtp"))
    pos, types, cell = configs.structure_c4(3000)
    ref = Oracle(path).compute(cell, types, pos, param_grad=True)
    return path, pos, types, cell, ref


@pytest.mark.parametrize("fp32", [False, True])
def test_c4_rational_efw_and_param_grad(c4_case, fp32):
    path, pos, types, cell, ref = c4_case
    print(f"[case] C4 N={len(types)} fp32={fp32}", flush=True)
    assert sus2sh.model_check(path)["species_count"] == 3
    pot = _engine(path, len(types), 3, fp32)
    got = _eval(pot, cell, types, pos, param_grad=True)
    pot.close()
    _check_efw(ref, got, fp32)
    g0, g1 = ref[3], got[3]
    err = np.abs(g1 - g0).max() / np.abs(g0).max()
    print(f"  dE/dtheta: {len(g0)} parameters, max rel err {err:.2e}", flush=True)
    assert err <= (1e-4 if fp32 else 1e-9)
    # the rational map's (a, s) adjoint: scal_coeffs are the last 2 S^2 Mr parameters
    nscal = 2 * 9 * 20
    assert np.abs(g0[-nscal:]).max() > 0
    assert np.abs(g1[-nscal:] - g0[-nscal:]).max() <= (1e-4 if fp32 else 1e-9) * np.abs(g0).max()


def test_c4_central_differences(c4_case):
    import torch
    path, pos, types, cell, ref = c4_case
    n = len(types)
    print(f"[case] C4 central differences N={n}", flush=True)
    pot = _engine(path, n, 3, False)
    e_a, f_a, _, g_a = _eval(pot, cell, types, pos, param_grad=True)
    t, x = _dev(types, np.int32), _dev(pos, np.float64)
    e = torch.zeros(n, dtype=torch.float64, device="cuda")
    f = torch.zeros(3, n, dtype=torch.float64, device="cuda")
    v = torch.zeros(9, n, dtype=torch.float64, device="cuda")
    box = sus2sh.box_struct(cell)

    def energy():
        e.zero_(), f.zero_(), v.zero_()
        pot.compute(box, t, x, e, f, v)
        torch.cuda.synchronize()
        return float(e.sum().item())

    h = 1e-4
    rng = np.random.default_rng(4004)
    for a in rng.choice(n, 6, replace=False):
        c = int(rng.integers(0, 3))
        x[c, a] += h
        ep = energy()
        x[c, a] -= 2 * h
        em = energy()
        x[c, a] += h
        fd = -(ep - em) / (2 * h)
        print(f"  atom {a} dir {c}: F={f_a[c, a]:.8f} fd={fd:.8f}", flush=True)
        assert abs(f_a[c, a] - fd) <= 1e-5
    theta = pot.get_params()
    P = len(theta)
    S, Q = 3, pot.info()["scalar_count"]
    # one parameter of each kind: shift, species, moment, radial, type, scal (a), scal (s)
    o_rad = 2 * S + Q
    picks = [0, S + 1, 2 * S + 3, o_rad + 7, P - 2 * 9 * 20 - 2, P - 2 * 9 * 20 + 5, P - 3,
             o_rad + 101]
    for k in picks:
        th = theta.copy()
        th[k] += h
        pot.set_params(th)
        ep = energy()
        th[k] -= 2 * h
        pot.set_params(th)
        em = energy()
        fd = (ep - em) / (2 * h)
        print(f"  theta[{k}]: dE={g_a[k]:.8e} fd={fd:.8e}", flush=True)
        assert abs(g_a[k] - fd) <= 1e-5 * max(1.0, abs(fd))
    pot.set_params(theta)
    pot.close()


@pytest.mark.parametrize("fp32", [False, True])
def test_c5_batch_efw_and_param_grad(model_dir, fp32):
    """C5-shaped batch (48 triclinic cells of 8..256 atoms from the C5 generator, most thinner
    than 2 rc, so multi-image neighbours) in one sus2sh_compute_batch call with dE/dtheta:
    per-atom E/F/W against the oracle cell by cell, and dE/dtheta against the sum of the
    oracle's per-cell gradients."""
    import torch
    path = sus2sh.write_mtp(configs.model_c5(), str(model_dir / "c5.mtp"))
    cells, pos, types, offs = configs.batch_c5(count=48, seed=5101)
    n = len(types)
    print(f"[case] C5 batch of {len(cells)} cells, N={n}, fp32={fp32}", flush=True)
    orc = Oracle(path)
    refs = [orc.compute(c, types[offs[k]:offs[k + 1]], pos[:, offs[k]:offs[k + 1]],
                        param_grad=True) for k, c in enumerate(cells)]
    pot = _engine(path, n, 2, fp32)
    t, x = _dev(types, np.int32), _dev(pos, np.float64)
    e = torch.zeros(n, dtype=torch.float64, device="cuda")
    f = torch.zeros(3, n, dtype=torch.float64, device="cuda")
    v = torch.zeros(9, n, dtype=torch.float64, device="cuda")
    g = torch.zeros(pot.info()["param_count"], dtype=torch.float64, device="cuda")
    sus2sh.compute_batch(pot, sus2sh.box_array(cells), offs, t, x, e, f, v, param_grad=g)
    torch.cuda.synchronize()
    pot.close()
    got = (e.cpu().numpy(), f.cpu().numpy(), v.cpu().numpy())
    ref = tuple(np.concatenate([r[i] for r in refs], axis=-1) for i in range(3))
    _check_efw(ref, got, fp32)
    g0 = np.sum([r[3] for r in refs], axis=0)
    err = np.abs(g.cpu().numpy() - g0).max() / np.abs(g0).max()
    print(f"  dE/dtheta: {len(g0)} parameters, max rel err {err:.2e}", flush=True)
    assert err <= (1e-4 if fp32 else 1e-9)
