"""Deterministic summation (sus2_sh_deterministic=1) on the GPU, through the C ABI.

The default force path sums neighbour forces with floating-point atomics, so its FP64 and
FP32 results depend on the order in which threads add them; the reference records the same
run-to-run scatter of its float path (docs/gpumd-sus2-sh-interface-plan.md:1099-1107).  With
the option, force contributions are summed in 64-bit fixed point (quantum 2^-34 eV/A) with
integer atomics, the graph kernel is chosen without timing, and row programs merge no back
chunk atomically, so energies, forces and virials are bit-identical between calls, handles
and MD runs.  Each case is also held to the oracle at the usual tolerances (FP64 1e-9
eV/atom, 1e-8 eV/A, 1e-8 eV; FP32 2e-5 eV/atom + 2e-6 |E|, 1e-4 eV/A + 2e-5 |F|max, 1e-4 eV
+ 2e-5 |W|max).

Cases: the benchmark graph (l3333 body 5, lane program) on 4,608 fcc Cu-Zr atoms (the size at
which the default path would autotune); a body-6 graph (no lane program: the row kernel);
a ZBL model on a dense random structure; a two-structure batch; 60 velocity-Verlet steps
(captured HIP graph, deferred checks) run twice; and the range check, forced from the host
with a dimer at 0.05 A whose ZBL pair force (~2e6 eV/A) exceeds the fixed-point bound.
"""
import numpy as np
import pytest

import sus2sh
from sus2sh import configs
from sus2sh import structures as st
from oracle.binding import Oracle

from test_gpu_parity import _compare, _separated_positions

pytestmark = pytest.mark.gpu


def _dev(a, dtype):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=dtype), device="cuda")


def _eval(pot, cell, types, pos, pbc=(1, 1, 1)):
    import torch
    n = len(types)
    t, x = _dev(types, np.int32), _dev(pos, np.float64)
    e = torch.zeros(n, dtype=torch.float64, device="cuda")
    f = torch.zeros(3, n, dtype=torch.float64, device="cuda")
    v = torch.zeros(9, n, dtype=torch.float64, device="cuda")
    pot.compute(sus2sh.box_struct(cell, pbc), t, x, e, f, v)
    torch.cuda.synchronize()
    return e.cpu().numpy(), f.cpu().numpy(), v.cpu().numpy()


def _identical(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def _repeat_case(path, cell, types, pos, fp32, symbols=("Cu", "Zr")):
    opts = list(symbols) + [f"sus2_float={int(fp32)}", "sus2_sh_deterministic=1"]
    n = len(types)
    pot = sus2sh.SUS2_SH(path, n, opts)
    first = _eval(pot, cell, types, pos)
    second = _eval(pot, cell, types, pos)
    info = pot.info()
    pot.close()
    pot = sus2sh.SUS2_SH(path, n, opts)
    third = _eval(pot, cell, types, pos)
    pot.close()
    assert _identical(first, second), "two calls of one handle differ"
    assert _identical(first, third), "two handles differ"
    return first, info


@pytest.mark.parametrize("fp32", [True, False])
def test_benchmark_graph_is_bit_reproducible(model_dir, fp32):
    spec = sus2sh.synthetic_model(3, 3, 5, body_l_max=(3, 3, 3, 3), species_count=2, rb_size=8,
                                  max_dist=5.0, seed=3003, scaling=0.5)
    path = sus2sh.write_mtp(spec, str(model_dir / "det_l3333.mtp"))
    pos, types, cell = st.fcc_alloy(12, 12, 8, 3.85, [0.5, 0.5], 0.08, 3001, 3002)
    print(f"[case] deterministic l3333 fp32={fp32} N={len(types)}", flush=True)
    got, info = _repeat_case(path, cell, types, pos, fp32)
    assert info["graph_variant"] == 3
    _compare(Oracle(path).compute(cell, types, pos), got, fp32)


@pytest.mark.parametrize("fp32", [True, False])
def test_body6_row_kernel_is_bit_reproducible(model_dir, fp32):
    spec = sus2sh.synthetic_model(2, 3, 6, species_count=2, rb_size=5, max_dist=4.2, seed=15,
                                  scaling=0.6)
    path = sus2sh.write_mtp(spec, str(model_dir / "det_body6.mtp"))
    pos, types, cell = st.fcc_alloy(6, 6, 6, 3.6, [0.5, 0.5], 0.05, 3011, 3012)
    print(f"[case] deterministic body-6 fp32={fp32} N={len(types)}", flush=True)
    got, info = _repeat_case(path, cell, types, pos, fp32)
    assert info["graph_variant"] != 3
    _compare(Oracle(path).compute(cell, types, pos), got, fp32)


def _zbl_model(model_dir):
    spec = sus2sh.synthetic_model(2, 2, 5, species_count=2, rb_size=8, max_dist=5.0, seed=41,
                                  scaling=0.05, pair_core=5.0)
    spec.extra = {"zbl_enabled": True, "zbl_inner": 0.8, "zbl_outer": 1.6}
    return sus2sh.write_mtp(spec, str(model_dir / "det_zbl.mtp"))


@pytest.mark.parametrize("fp32", [True, False])
def test_zbl_is_bit_reproducible(model_dir, fp32):
    path = _zbl_model(model_dir)
    rng = np.random.default_rng(5)
    n = 120
    L = (n / 0.07) ** (1 / 3)
    cell = np.diag([L, L, L])
    pos = _separated_positions(cell, n, 0.95, rng)
    types = rng.integers(0, 2, size=n).astype(np.int32)
    print(f"[case] deterministic ZBL fp32={fp32} N={n}", flush=True)
    got, _ = _repeat_case(path, cell, types, pos, fp32)
    z, inner, outer = sus2sh.zbl_tables(path, ["Cu", "Zr"])
    orc = Oracle(path)
    orc.set_zbl(z, inner, outer)
    _compare(orc.compute(cell, types, pos), got, fp32)


def test_batch_is_bit_reproducible(model_dir):
    import torch
    path = sus2sh.write_mtp(configs.model_c2(md=False), str(model_dir / "det_batch.mtp"))
    structs = [st.fcc_alloy(4, 4, 4, 3.85, [0.64, 0.36], 0.08, 3021, 3022),
               st.fcc_alloy(5, 4, 4, 3.85, [0.64, 0.36], 0.08, 3023, 3024)]
    types = np.concatenate([s[1] for s in structs]).astype(np.int32)
    pos = np.concatenate([s[0] for s in structs], axis=1)
    offs = np.cumsum([0] + [len(s[1]) for s in structs])
    n = len(types)
    print(f"[case] deterministic batch of 2, N={n}", flush=True)
    pot = sus2sh.SUS2_SH(path, n, ["Cu", "Zr", "sus2_float=1", "sus2_sh_deterministic=1"])
    outs = []
    for _ in range(2):
        t, x = _dev(types, np.int32), _dev(pos, np.float64)
        e = torch.zeros(n, dtype=torch.float64, device="cuda")
        f = torch.zeros(3, n, dtype=torch.float64, device="cuda")
        v = torch.zeros(9, n, dtype=torch.float64, device="cuda")
        sus2sh.compute_batch(pot, [s[2] for s in structs], offs, t, x, e, f, v)
        torch.cuda.synchronize()
        outs.append((e.cpu().numpy(), f.cpu().numpy(), v.cpu().numpy()))
    pot.close()
    assert _identical(outs[0], outs[1])
    orc = Oracle(path)
    for k, (p, ty, c) in enumerate(structs):
        a, b = offs[k], offs[k + 1]
        _compare(orc.compute(c, ty, p), tuple(o[..., a:b] for o in outs[0]), True)


@pytest.mark.parametrize("fp32", [True, False])
def test_md_trajectories_are_bit_identical(model_dir, fp32):
    import torch
    path = sus2sh.write_mtp(configs.model_c2(md=True), str(model_dir / "det_md.mtp"))
    pos, types, cell = configs.structure_c2(6)
    masses = np.array([configs.MASSES["Cu"], configs.MASSES["Zr"]])
    vel = sus2sh.maxwell_velocities(masses[types], 300.0, 1007)
    n = len(types)
    print(f"[case] deterministic NVE fp32={fp32} N={n} 60 steps x 2", flush=True)
    runs = []
    for _ in range(2):
        pot = sus2sh.SUS2_SH(path, n, ["Cu", "Zr", f"sus2_float={int(fp32)}",
                                       "sus2_sh_deterministic=1"])
        t, x, v = _dev(types, np.int32), _dev(pos, np.float64), _dev(vel, np.float64)
        trace, e, f, w = sus2sh.md_nve(pot, sus2sh.box_struct(cell), t, x, v, masses, 0.5, 60)
        torch.cuda.synchronize()
        pot.close()
        runs.append((x.cpu().numpy(), v.cpu().numpy(), f.cpu().numpy(), trace))
    for name, a, b in zip(("positions", "velocities", "forces", "energy trace"), runs[0],
                          runs[1]):
        assert np.array_equal(a, b), f"{name} differ between the two runs"


def test_out_of_range_pair_force_is_reported(model_dir):
    path = _zbl_model(model_dir)
    cell = np.diag([14.0, 14.0, 14.0])
    pos = np.array([[5.0, 5.05], [5.0, 5.0], [5.0, 5.0]])
    types = np.array([0, 1], dtype=np.int32)
    print("[case] deterministic range check: ZBL dimer at 0.05 A", flush=True)
    pot = sus2sh.SUS2_SH(path, 2, ["Cu", "Zr", "sus2_float=0", "sus2_sh_deterministic=1"])
    with pytest.raises(sus2sh.Sus2shError, match="fixed-point range"):
        _eval(pot, cell, types, pos)
    # the flag is cleared: a structure in range evaluates again on the same handle
    pos_ok = np.array([[5.0, 7.4], [5.0, 5.0], [5.0, 5.0]])
    e, f, _ = _eval(pot, cell, types, pos_ok)
    pot.close()
    assert np.all(np.isfinite(e)) and np.abs(f[:, 0] + f[:, 1]).max() < 1e-9


def test_param_grad_is_rejected(model_dir):
    import torch
    path = sus2sh.write_mtp(configs.model_c2(md=False), str(model_dir / "det_pg.mtp"))
    pos, types, cell = st.fcc_alloy(3, 3, 3, 3.85, [0.5, 0.5], 0.05, 1, 2)
    n = len(types)
    pot = sus2sh.SUS2_SH(path, n, ["Cu", "Zr", "sus2_sh_deterministic=1"])
    t, x = _dev(types, np.int32), _dev(pos, np.float64)
    e = torch.zeros(n, dtype=torch.float64, device="cuda")
    f = torch.zeros(3, n, dtype=torch.float64, device="cuda")
    v = torch.zeros(9, n, dtype=torch.float64, device="cuda")
    g = torch.zeros(pot.info()["param_count"], dtype=torch.float64, device="cuda")
    with pytest.raises(sus2sh.Sus2shError, match="deterministic"):
        pot.compute(sus2sh.box_struct(cell), t, x, e, f, v, param_grad=g)
    pot.close()
