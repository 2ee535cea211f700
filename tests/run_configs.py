#!/usr/bin/env python3
"""Run the parity/benchmark configurations C1, C2, C4, C5 (C3 is bench.py) and the
reference-shaped C3 variant on one GPU and print one JSON record per configuration.

  python tests/run_configs.py [--configs c1 c2 c4 c5 c3ref] [--quick]

Errors are reported against the engine's FP64 path and, where the FP64 CPU oracle finishes
in seconds, against the oracle.  --quick shrinks every workload for a smoke run.

c3ref is the C3 structure (80x80x40 fcc Cu-Zr, 1,024,000 atoms) under a model of the shape
the reference's A100 figures are quoted on (body-5 l3333, rb = 10, rc = 7.5 A,
docs/gpumd-sus2-sh-interface-plan.md:24-40 and :1409), timed as a fixed-cell force loop.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "gpumd-sus2-sh_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _write(spec, name):
    import sus2sh
    d = tempfile.mkdtemp(prefix="sus2sh_cfg_")
    return sus2sh.write_mtp(spec, os.path.join(d, name + ".mtp"))


def _dev(a, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a) if dtype is None else
                           np.ascontiguousarray(a, dtype=dtype), device="cuda")


def _eval(path, cell, types, pos, fp32, param_grad=False):
    import torch
    import sus2sh
    n = len(types)
    pot = sus2sh.SUS2_SH(path, n, ["A", "B", "C"][: 2 if types.max() < 2 else 3] +
                         [f"sus2_float={int(fp32)}"])
    t, x = _dev(types, np.int32), _dev(pos, np.float64)
    e = torch.zeros(n, dtype=torch.float64, device="cuda")
    f = torch.zeros(3, n, dtype=torch.float64, device="cuda")
    v = torch.zeros(9, n, dtype=torch.float64, device="cuda")
    g = torch.zeros(pot.info()["param_count"], dtype=torch.float64, device="cuda") \
        if param_grad else None
    box = sus2sh.box_struct(cell)
    pot.compute(box, t, x, e, f, v, param_grad=g)  # first call (autotune, allocation)
    torch.cuda.synchronize()
    e.zero_(), f.zero_(), v.zero_()
    if g is not None:
        g.zero_()
    t0 = time.perf_counter()
    pot.compute(box, t, x, e, f, v, param_grad=g)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out = dict(e=e.cpu().numpy(), f=f.cpu().numpy(), v=v.cpu().numpy(), seconds=dt,
               g=None if g is None else g.cpu().numpy(), nbar=pot.mean_neighbors(),
               info=pot.info())
    pot.close()
    return out


def _err(a, b):
    return dict(energy_per_atom=float(np.abs(a["e"] - b[0]).max()),
                force=float(np.abs(a["f"] - b[1]).max()),
                virial=float(np.abs(a["v"] - b[2]).max()))


def _nve(path, cell, types, pos, masses, dt, steps, fp32):
    import torch
    import sus2sh
    n = len(types)
    pot = sus2sh.SUS2_SH(path, n, ["Cu", "Zr", f"sus2_float={int(fp32)}"])
    vel = sus2sh.maxwell_velocities(masses[types], 300.0, 1004)
    t, x, vv = _dev(types, np.int32), _dev(pos, np.float64), _dev(vel, np.float64)
    t0 = time.perf_counter()
    trace, *_ = sus2sh.md_nve(pot, sus2sh.box_struct(cell), t, x, vv, masses, dt, steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    pot.close()
    tot = trace.sum(axis=1)
    kB = 8.617333262e-5
    return dict(steps=steps, dt_fs=dt, seconds=el, atom_steps_per_s=n * steps / el,
                drift_ev_per_atom=float((tot[-1] - tot[0]) / n),
                max_dev_ev_per_atom=float(np.abs(tot - tot[0]).max() / n),
                drift_ev_per_atom_per_ps=float((tot[-1] - tot[0]) / n / (steps * dt * 1e-3)),
                t_final_K=float(trace[-1, 1] / (1.5 * n * kB))), x.cpu().numpy()


def run_c1(quick):
    from sus2sh import configs
    from oracle.binding import Oracle
    spec = configs.model_c1()
    path = _write(spec, "c1_l2k2")
    pos, types, cell = configs.structure_c1(5 if quick else 10)
    ref = Oracle(path).compute(cell, types, pos)
    r64, r32 = _eval(path, cell, types, pos, False), _eval(path, cell, types, pos, True)
    masses = np.array([configs.MASSES["Cu"], configs.MASSES["Zr"]])
    steps = 20 if quick else 100
    md64, x64 = _nve(path, cell, types, pos, masses, 1.0, steps, False)
    md32, x32 = _nve(path, cell, types, pos, masses, 1.0, steps, True)
    return dict(config="C1", atoms=len(types), model="l2k2 body5 rb8 rc5.0 seed1003 pair_core5",
                fp64_vs_oracle=_err(r64, ref), fp32_vs_oracle=_err(r32, ref),
                nve_fp64=md64, nve_fp32=md32,
                nve_fp32_vs_fp64_max_position_dev=float(np.abs(x32 - x64).max()))


def run_c2(quick):
    from sus2sh import configs
    from oracle.binding import Oracle
    spec = configs.model_c2(md=True)
    path = _write(spec, "c2_l3k3")
    pos, types, cell = configs.structure_c2(8 if quick else 20)
    ref = Oracle(path).compute(cell, types, pos)
    r64, r32 = _eval(path, cell, types, pos, False), _eval(path, cell, types, pos, True)
    masses = np.array([configs.MASSES["Cu"], configs.MASSES["Zr"]])
    steps = 50 if quick else 1000
    md64, _ = _nve(path, cell, types, pos, masses, 0.5, steps, False)
    md32, _ = _nve(path, cell, types, pos, masses, 0.5, steps, True)
    return dict(config="C2", atoms=len(types), model="l3k3 body5 rb8 rc5.0 seed2003 pair_core5",
                fp64_vs_oracle=_err(r64, ref), fp32_vs_oracle=_err(r32, ref),
                fp32_vs_fp64=_err(r32, (r64["e"], r64["f"], r64["v"])),
                nve_fp64=md64, nve_fp32=md32)


def run_c4(quick):
    from sus2sh import configs
    from oracle.binding import Oracle
    spec = configs.model_c4()
    path = _write(spec, "c4_l4k4")
    n = 8000 if quick else 256000
    t0 = time.perf_counter()
    pos, types, cell = configs.structure_c4(n)
    log(f"[c4] structure {n} atoms in {time.perf_counter() - t0:.1f} s")
    r32 = _eval(path, cell, types, pos, True, param_grad=True)
    r64 = _eval(path, cell, types, pos, False, param_grad=True)
    out = dict(config="C4", atoms=n, model="l4k4 body5 rb10 rational rc5.5 S=3 seed4003",
               graph_moments=r32["info"]["moments_count"],
               graph_products=r32["info"]["product_count"], mean_neighbors=r32["nbar"],
               fp32_seconds=r32["seconds"], fp64_seconds=r64["seconds"],
               fp32_atom_evals_per_s=n / r32["seconds"],
               fp32_vs_fp64=_err(r32, (r64["e"], r64["f"], r64["v"])),
               param_grad_fp32_vs_fp64_rel=float(np.abs(r32["g"] - r64["g"]).max() /
                                                 max(1.0, np.abs(r64["g"]).max())))
    out.update(_c4_central_differences(path, cell, types, pos, r32, r64))
    if quick or n <= 64000:
        ref = Oracle(path).compute(cell, types, pos, param_grad=True)
        out["fp64_vs_oracle"] = _err(r64, ref)
        out["param_grad_fp64_vs_oracle_rel"] = float(np.abs(r64["g"] - ref[3]).max() /
                                                     max(1.0, np.abs(ref[3]).max()))
    return out


def _c4_central_differences(path, cell, types, pos, r32, r64, h=1e-4, samples=6):
    """Central differences (h = 1e-4) of the FP64 engine's total energy on a subsample of
    atom coordinates and of model parameters, against the analytic FP64 and FP32 forces and
    dE/dtheta."""
    import torch
    import sus2sh
    n = len(types)
    rng = np.random.default_rng(4004)
    pot = sus2sh.SUS2_SH(path, n, ["A", "B", "C", "sus2_float=0"])
    t, x = _dev(types, np.int32), _dev(pos, np.float64)
    e = torch.zeros(n, dtype=torch.float64, device="cuda")
    f = torch.zeros(3, n, dtype=torch.float64, device="cuda")
    v = torch.zeros(9, n, dtype=torch.float64, device="cuda")
    box = sus2sh.box_struct(cell)

    def energy():
        e.zero_(), f.zero_(), v.zero_()
        pot.compute(box, t, x, e, f, v)
        return float(e.sum().item())

    fd_f = []
    for a in rng.choice(n, samples, replace=False):
        c = int(rng.integers(0, 3))
        x[c, a] += h
        ep = energy()
        x[c, a] -= 2 * h
        em = energy()
        x[c, a] += h
        fd = -(ep - em) / (2 * h)
        fd_f.append((abs(r64["f"][c, a] - fd), abs(r32["f"][c, a] - fd)))
    theta = pot.get_params()
    fd_g = []
    for k in rng.choice(len(theta), samples, replace=False):
        th = theta.copy()
        th[k] += h
        pot.set_params(th)
        ep = energy()
        th[k] -= 2 * h
        pot.set_params(th)
        em = energy()
        fd = (ep - em) / (2 * h)
        scale = max(1.0, abs(fd))
        fd_g.append((abs(r64["g"][k] - fd) / scale, abs(r32["g"][k] - fd) / scale))
    pot.set_params(theta)
    pot.close()
    fd_f, fd_g = np.array(fd_f), np.array(fd_g)
    return dict(fd_samples=samples, fd_step=h,
                fd_force_err_fp64=float(fd_f[:, 0].max()), fd_force_err_fp32=float(fd_f[:, 1].max()),
                fd_param_grad_rel_err_fp64=float(fd_g[:, 0].max()),
                fd_param_grad_rel_err_fp32=float(fd_g[:, 1].max()))


def run_c5(quick):
    import torch
    import sus2sh
    from sus2sh import configs
    from oracle.binding import Oracle
    spec = configs.model_c5()
    path = _write(spec, "c5_l3k3")
    count = 256 if quick else 4096
    cells, pos, types, offs = configs.batch_c5(count)
    N = len(types)
    res = {}
    for fp32 in (True, False):
        pot = sus2sh.SUS2_SH(path, N, ["Cu", "Zr", f"sus2_float={int(fp32)}"])
        P = pot.info()["param_count"]
        t, x = _dev(types, np.int32), _dev(pos, np.float64)
        e = torch.zeros(N, dtype=torch.float64, device="cuda")
        f = torch.zeros(3, N, dtype=torch.float64, device="cuda")
        v = torch.zeros(9, N, dtype=torch.float64, device="cuda")
        g = torch.zeros(P, dtype=torch.float64, device="cuda")
        sus2sh.compute_batch(pot, cells, offs, t, x, e, f, v, param_grad=g)
        torch.cuda.synchronize()
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            e.zero_(), f.zero_(), v.zero_(), g.zero_()
            sus2sh.compute_batch(pot, cells, offs, t, x, e, f, v, param_grad=g)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        res[fp32] = dict(seconds=dt, e=e.cpu().numpy(), f=f.cpu().numpy(), v=v.cpu().numpy(),
                         g=g.cpu().numpy())
        pot.close()
    orc = Oracle(path)
    errs = []
    for k in range(0, count, max(1, count // 16)):
        a, b = offs[k], offs[k + 1]
        e0, f0, v0 = orc.compute(cells[k], types[a:b], pos[:, a:b], brute=True)
        errs.append((np.abs(res[False]["e"][a:b] - e0).max(),
                     np.abs(res[False]["f"][:, a:b] - f0).max(),
                     np.abs(res[True]["f"][:, a:b] - f0).max()))
    errs = np.array(errs)
    r32 = res[True]
    return dict(config="C5", structures=count, atoms=int(N),
                fp32_seconds=r32["seconds"], structures_per_s=count / r32["seconds"],
                atom_evals_per_s=N / r32["seconds"], fp64_seconds=res[False]["seconds"],
                with_param_grad=True,
                fp32_vs_fp64_force=float(np.abs(r32["f"] - res[False]["f"]).max()),
                param_grad_fp32_vs_fp64_rel=float(np.abs(r32["g"] - res[False]["g"]).max() /
                                                  max(1.0, np.abs(res[False]["g"]).max())),
                oracle_sample_structures=len(errs),
                fp64_vs_oracle_energy=float(errs[:, 0].max()),
                fp64_vs_oracle_force=float(errs[:, 1].max()),
                fp32_vs_oracle_force=float(errs[:, 2].max()))


def run_c3ref(quick):
    import torch
    import sus2sh
    from sus2sh import structures as st
    spec = sus2sh.synthetic_model(3, 3, 5, body_l_max=(3, 3, 3, 3), species_count=2,
                                  rb_size=10, max_dist=7.5, seed=3003, scaling=0.5)
    path = _write(spec, "c3ref_l3333_rb10_rc75")
    nx, ny, nz = (20, 20, 10) if quick else (80, 80, 40)
    pos, types, cell = st.fcc_alloy(nx, ny, nz, 3.85, [0.5, 0.5], 0.08, 3001, 3002)
    n = len(types)
    out = dict(config="C3ref", atoms=n, model="l3333 body5 rb10 rc7.5 seed3003",
               reference_a100_atom_steps_per_s=1.19500e7)
    forces = {}
    for fp32, steps in ((True, 5 if quick else 20), (False, 2 if quick else 5)):
        pot = sus2sh.SUS2_SH(path, n, ["Cu", "Zr", f"sus2_float={int(fp32)}"])
        t, x = _dev(types, np.int32), _dev(pos, np.float64)
        e = torch.zeros(n, dtype=torch.float64, device="cuda")
        f = torch.zeros(3, n, dtype=torch.float64, device="cuda")
        v = torch.zeros(9, n, dtype=torch.float64, device="cuda")
        box = sus2sh.box_struct(cell)
        for _ in range(2):
            e.zero_(), f.zero_(), v.zero_()
            pot.compute(box, t, x, e, f, v)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            e.zero_(), f.zero_(), v.zero_()
            pot.compute(box, t, x, e, f, v)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        key = "fp32" if fp32 else "fp64"
        out[key + "_ms_per_step"] = 1e3 * dt
        out[key + "_atom_steps_per_s"] = n / dt
        out["mean_neighbors"] = pot.mean_neighbors()
        forces[key] = f.cpu().numpy()
        pot.close()
    out["fp32_over_fp64"] = out["fp32_atom_steps_per_s"] / out["fp64_atom_steps_per_s"]
    out["max_force_err_fp32_vs_fp64"] = float(np.abs(forces["fp32"] - forces["fp64"]).max())
    out["rms_force_err_fp32_vs_fp64"] = float(np.sqrt(((forces["fp32"] - forces["fp64"]) ** 2).mean()))
    out["max_abs_force_fp64"] = float(np.abs(forces["fp64"]).max())
    out["rms_force_fp64"] = float(np.sqrt((forces["fp64"] ** 2).mean()))
    out["vs_reference_a100"] = out["fp32_atom_steps_per_s"] / 1.19500e7
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", nargs="+", default=["c1", "c2", "c4", "c5", "c3ref"])
    ap.add_argument("--quick", action="store_true")
    args = ap.parse_args()
    import torch
    assert torch.cuda.is_available(), "run_configs.py needs a HIP device"
    runners = dict(c1=run_c1, c2=run_c2, c4=run_c4, c5=run_c5, c3ref=run_c3ref)
    for c in args.configs:
        log(f"[config] {c} start")
        t0 = time.perf_counter()
        rec = runners[c](args.quick)
        rec["wall_seconds"] = time.perf_counter() - t0
        print(json.dumps(rec), flush=True)
        log(f"[config] {c} done in {rec['wall_seconds']:.1f} s")


if __name__ == "__main__":
    main()
