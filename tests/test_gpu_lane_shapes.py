"""GPU parity of every lane-kernel shape on the benchmark model (through the C ABI).

The FP32 benchmark runs the shape the autotuner picks for the l3333 graph (body order 5,
sh_body_l_max = {3,3,3,3}: 1349 moments, 5475 products, 572 scalars) among the twelve lane
shapes (waves per tile, accumulators per wave, slot budget, atoms per lane).  Each shape is pinned
with sus2_sh_lane_shape and compared with the FP64 oracle on a 12x12x8-cell fcc Cu-Zr piece
of the benchmark structure (4,608 atoms, a = 3.85 A, jitter 0.08 A, rc 5.0 A, rb 8,
seed 3003) -- the graph semantics of src/force/sus2_sh.cu:4799-4846.  A last case leaves the
autotuner on (it runs from 4,096 atoms) and checks the shape it keeps.

Tolerances as in test_gpu_parity: FP64 1e-9 eV/atom, 1e-8 eV/A, 1e-8 eV;
FP32 2e-5 eV/atom + 2e-6 |E|, 1e-4 eV/A + 2e-5 |F|max, 1e-4 eV + 2e-5 |W|max.
"""
import numpy as np
import pytest

import sus2sh
from sus2sh import structures as st
from oracle.binding import Oracle

pytestmark = pytest.mark.gpu

SHAPES = [(4, 4, 160, 1), (8, 2, 160, 1), (4, 2, 160, 1), (8, 2, 192, 1), (8, 2, 256, 1),
          (2, 4, 160, 1), (8, 2, 208, 1), (8, 2, 160, 2), (4, 2, 160, 2), (8, 2, 208, 2),
          (16, 2, 192, 1), (16, 2, 320, 1)]


@pytest.fixture(scope="module")
def bench_case(model_dir):
    spec = sus2sh.synthetic_model(3, 3, 5, body_l_max=(3, 3, 3, 3), species_count=2, rb_size=8,
                                  max_dist=5.0, seed=3003, scaling=0.5)
    path = sus2sh.write_mtp(spec, str(model_dir / "bench_l3333.mtp"))
    pos, types, cell = st.fcc_alloy(12, 12, 8, 3.85, [0.5, 0.5], 0.08, 3001, 3002)
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
    info = pot.info()
    out = (e.numpy(), f.numpy(), v.numpy())
    pot.close()
    return out, info


@pytest.mark.parametrize("fp32", [True, False])
@pytest.mark.parametrize("waves,nslot,slots,atoms", SHAPES)
def test_lane_shape_matches_oracle(bench_case, waves, nslot, slots, atoms, fp32):
    path, pos, types, cell, ref = bench_case
    if slots * 64 * atoms * (4 if fp32 else 8) > 160 * 1024:
        pytest.skip("the tile's LDS (slots x 64 x atoms x value bytes) exceeds 160 KiB")
    opts = ["Cu", "Zr", f"sus2_float={int(fp32)}", "sus2_sh_graph_variant=3",
            f"sus2_sh_lane_shape={waves}x{nslot}x{slots}x{atoms}"]
    print(f"[case] lane shape ({waves},{nslot},{slots},{atoms}) fp32={fp32} N={len(types)}",
          flush=True)
    got, info = _evaluate(path, cell, types, pos, opts)
    assert info["graph_variant"] == 3
    assert (info["lane_waves"], info["lane_nslot"], info["lane_atoms"]) == (waves, nslot, atoms)
    assert info["lane_slots"] <= slots
    _compare(ref, got, fp32)


@pytest.mark.parametrize("fp32", [True, False])
def test_autotuned_shape_matches_oracle(bench_case, fp32):
    """N = 4,608 >= 4,096: the first call times every graph-kernel candidate and keeps the
    fastest; its outputs (from the kept candidate's launch) are checked against the oracle."""
    path, pos, types, cell, ref = bench_case
    print(f"[case] autotuned fp32={fp32} N={len(types)}", flush=True)
    got, info = _evaluate(path, cell, types, pos, ["Cu", "Zr", f"sus2_float={int(fp32)}"])
    print(f"  kept variant {info['graph_variant']} tile {info['graph_tile']} lane shape "
          f"({info['lane_waves']},{info['lane_nslot']},{info['lane_slots']},"
          f"{info['lane_atoms']})", flush=True)
    _compare(ref, got, fp32)
