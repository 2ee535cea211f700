"""GPU parity of every edge-kernel launch shape (through the C ABI).

The basic-moment and force kernels run with `threads` per workgroup (64, 128, 256) and `npl`
lanes per atom (1, 2, 4, 8): lane s of an atom takes its edges s, s + npl, ... and the npl
partial (k, l, m) grids, forces and virials are summed with lane shuffles.  Each shape is
pinned with sus2_sh_basic_shape / sus2_sh_force_shape and compared with the FP64 oracle on a
2,048-atom periodic fcc Cu-Zr box (8 force tiles of 256 atoms whose neighbour forces are
summed in LDS windows: a workgroup of threads / npl atoms uses its tile's windows), in both
precisions; a second case checks a batch (one window per tile: the sorted range of its
structures) and the autotuner, which
times all 12 x 2 candidates on the first call from 512 atoms on.

Tolerances as in test_gpu_parity: FP64 1e-9 eV/atom, 1e-8 eV/A, 1e-8 eV;
FP32 2e-5 eV/atom + 2e-6 |E|, 1e-4 eV/A + 2e-5 |F|max, 1e-4 eV + 2e-5 |W|max.
"""
import numpy as np
import pytest

import sus2sh
from sus2sh import structures as st
from oracle.binding import Oracle

pytestmark = pytest.mark.gpu

SHAPES = [(t, n) for t in (256, 128, 64) for n in (1, 2, 4, 8)]


@pytest.fixture(scope="module")
def case(model_dir):
    spec = sus2sh.synthetic_model(3, 3, 5, species_count=2, rb_size=8, max_dist=5.0, seed=12,
                                  scaling=0.5)
    path = sus2sh.write_mtp(spec, str(model_dir / "edge_l3k3.mtp"))
    pos, types, cell = st.fcc_alloy(8, 8, 8, 3.85, [0.5, 0.5], 0.08, 31, 32)
    ref = Oracle(path).compute(cell, types, pos)
    return path, pos, types, cell, ref


def _compare(ref, got, fp32):
    e0, f0, v0 = ref
    e1, f1, v1 = got
    de, df, dv = np.abs(e1 - e0).max(), np.abs(f1 - f0).max(), np.abs(v1 - v0).max()
    print(f"  max|dE|={de:.3e} max|dF|={df:.3e} max|dW|={dv:.3e}", flush=True)
    if fp32:
        assert de <= 2e-5 + 2e-6 * np.abs(e0).max()
        assert df <= 1e-4 + 2e-5 * np.abs(f0).max()
        assert dv <= 1e-4 + 2e-5 * np.abs(v0).max()
    else:
        assert de <= 1e-9 and df <= 1e-8 and dv <= 1e-8


def _evaluate(path, cell, types, pos, opts):
    from sus2sh.hipmem import DeviceArray, synchronize
    n = len(types)
    pot = sus2sh.SUS2_SH(path, n, opts)
    t = DeviceArray.from_numpy(np.asarray(types, dtype=np.int32))
    x = DeviceArray.from_numpy(np.ascontiguousarray(pos, dtype=np.float64))
    e = DeviceArray.zeros((n,), np.float64)
    f = DeviceArray.zeros((3, n), np.float64)
    v = DeviceArray.zeros((9, n), np.float64)
    pot.compute(sus2sh.box_struct(cell), t, x, e, f, v)
    synchronize()
    shapes = pot.edge_shapes()
    out = (e.numpy(), f.numpy(), v.numpy())
    pot.close()
    return out, shapes


@pytest.mark.parametrize("fp32", [True, False])
@pytest.mark.parametrize("threads,npl", SHAPES)
def test_edge_shape_matches_oracle(case, threads, npl, fp32):
    path, pos, types, cell, ref = case
    opts = ["Cu", "Zr", f"sus2_float={int(fp32)}", f"sus2_sh_basic_shape={threads}x{npl}",
            f"sus2_sh_force_shape={threads}x{npl}"]
    print(f"[case] edge shape {threads}x{npl} fp32={fp32} N={len(types)}", flush=True)
    got, (shape, times) = _evaluate(path, cell, types, pos, opts)
    assert shape == {"basic": (threads, npl), "force": (threads, npl), "basic_mfma": 0}
    assert times == {}  # fixed shapes: the edge autotuner does not run
    _compare(ref, got, fp32)


@pytest.mark.parametrize("fp32", [True, False])
def test_edge_autotuner_and_batch(case, fp32):
    """The first call (N = 2,048 >= 512) times all 24 candidates and keeps the fastest shape
    of each kernel; its outputs are checked against the oracle.  Then the same structure
    twice in one batched call (no LDS windows) with 4 lanes per atom."""
    path, pos, types, cell, ref = case
    print(f"[case] edge autotuner fp32={fp32} N={len(types)}", flush=True)
    got, (shape, times) = _evaluate(path, cell, types, pos, ["Cu", "Zr", f"sus2_float={int(fp32)}"])
    print(f"  kept {shape}; times {times}", flush=True)
    # 12 vector basic-kernel shapes, 12 force shapes and, in FP32, the matrix-core basic kernel
    assert len(times) == (25 if fp32 else 24) and all(ms > 0 for ms in times.values())
    best_basic = min((ms, k) for k, ms in times.items() if k[0] in ("basic", "basic_mfma_fp32"))[1]
    best_force = min((ms, k) for k, ms in times.items() if k[0] == "force")[1]
    if best_basic[0] == "basic_mfma_fp32":
        assert shape["basic_mfma"] == 1
    else:
        assert shape["basic_mfma"] == 0 and shape["basic"] == best_basic[1:]
    assert shape["force"] == best_force[1:]
    _compare(ref, got, fp32)

    import torch
    n = len(types)
    opts = ["Cu", "Zr", f"sus2_float={int(fp32)}", "sus2_sh_basic_shape=128x4",
            "sus2_sh_force_shape=256x4"]
    pot = sus2sh.SUS2_SH(path, 2 * n, opts)
    dev = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a, dtype=dt), device="cuda")
    e = torch.zeros(2 * n, dtype=torch.float64, device="cuda")
    f = torch.zeros(3, 2 * n, dtype=torch.float64, device="cuda")
    v = torch.zeros(9, 2 * n, dtype=torch.float64, device="cuda")
    sus2sh.compute_batch(pot, [cell, cell], np.array([0, n, 2 * n], dtype=np.int32),
                         dev(np.concatenate([types, types]), np.int32),
                         dev(np.concatenate([pos, pos], axis=1), np.float64), e, f, v)
    torch.cuda.synchronize()
    pot.close()
    e, f, v = e.cpu().numpy(), f.cpu().numpy(), v.cpu().numpy()
    for k in range(2):
        sl = slice(k * n, (k + 1) * n)
        _compare(ref, (e[sl], f[:, sl], v[:, sl]), fp32)
