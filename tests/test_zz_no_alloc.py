"""GPU: the allocation-free entry point sus2sh_compute_noalloc (work buffers in the fixed
per-device scratch array) against sus2sh_compute on the same handle, through the C ABI.

Two calls in a row per precision, the second with a smaller input (so the scratch array is
re-carved and reused with the first call's contents still in it), each compared with
sus2sh_compute on identical device inputs.  Both paths run the same kernels; neighbour forces
are summed with floating-point atomics, so the bar is the summation-order bound: FP64
1e-12 relative to the largest |value| (energies, forces, virials), FP32 2e-6.  A last case
asks for more atoms than the 128 MiB scratch holds and expects SUS2SH_ERR_ARGUMENT.
"""
import ctypes

import numpy as np
import pytest

import sus2sh
from sus2sh import structures as st
from sus2sh._capi import lib
from sus2sh.hipmem import DeviceArray, synchronize

pytestmark = pytest.mark.gpu


def _arrays(pos, types):
    n = len(types)
    return dict(t=DeviceArray.from_numpy(np.asarray(types, dtype=np.int32)),
                x=DeviceArray.from_numpy(np.ascontiguousarray(pos, dtype=np.float64)),
                e=DeviceArray.zeros((n,), np.float64), f=DeviceArray.zeros((3, n), np.float64),
                v=DeviceArray.zeros((9, n), np.float64))


def _call(fn, pot, cell, a, n):
    box = sus2sh.box_struct(cell)
    p = lambda d: ctypes.c_void_p(d.data_ptr())  # noqa: E731
    return fn(pot._h, ctypes.byref(box), n, p(a["t"]), p(a["x"]), p(a["e"]), p(a["f"]),
              p(a["v"]), ctypes.c_void_p(0))


@pytest.mark.parametrize("fp32", [False, True])
def test_noalloc_matches_compute(make_model, fp32):
    path = make_model("noalloc_l3k3", l_max=3, k_max=3, body_order=5, body_l_max=(3, 3, 3, 3),
                      rb_size=8, max_dist=5.0, seed=3003, scaling=0.5)
    L = lib()
    pot = sus2sh.SUS2_SH(path, 0, ["Cu", "Zr", f"sus2_float={int(fp32)}"])
    tol = 2e-6 if fp32 else 1e-12
    for cells in ((10, 10, 8), (6, 5, 4)):  # 3,200 atoms, then 480
        pos, types, cell = st.fcc_alloy(*cells, 3.85, [0.5, 0.5], 0.08, 3001, 3002)
        n = len(types)
        print(f"[case] noalloc fp32={fp32} N={n}", flush=True)
        ref, got = _arrays(pos, types), _arrays(pos, types)
        assert _call(L.sus2sh_compute, pot, cell, ref, n) == 0, sus2sh.last_error(pot._h)
        rc = _call(L.sus2sh_compute_noalloc, pot, cell, got, n)
        assert rc == 0, sus2sh.last_error(pot._h)
        synchronize()
        for k in ("e", "f", "v"):
            a, b = ref[k].numpy(), got[k].numpy()
            scale = max(1.0, np.abs(a).max())
            err = np.abs(a - b).max() / scale
            print(f"  {k}: max rel diff {err:.2e}", flush=True)
            assert err <= tol
    pot.close()


def test_noalloc_rejects_inputs_beyond_the_scratch(make_model):
    path = make_model("noalloc_l3k3", l_max=3, k_max=3, body_order=5, body_l_max=(3, 3, 3, 3),
                      rb_size=8, max_dist=5.0, seed=3003, scaling=0.5)
    pot = sus2sh.SUS2_SH(path, 0, ["Cu", "Zr", "sus2_float=0"])
    # 186,624 atoms: the FP64 per-atom buffers alone (~1 KB/atom) exceed 128 MiB
    pos, types, cell = st.fcc_alloy(36, 36, 36, 3.85, [0.5, 0.5], 0.08, 3001, 3002)
    n = len(types)
    print(f"[case] noalloc oversize N={n}", flush=True)
    a = _arrays(pos, types)
    rc = _call(lib().sus2sh_compute_noalloc, pot, cell, a, n)
    msg = sus2sh.last_error(pot._h)
    print(f"  rc={rc}: {msg}", flush=True)
    assert rc == 2 and "scratch" in msg
    pot.close()
