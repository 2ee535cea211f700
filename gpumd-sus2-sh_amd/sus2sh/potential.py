"""Python mirror of the GPUMD plugin interface, on top of the C ABI.

`SUS2_SH(file, num_atoms, options)` and `SUS2_SH.compute(box, type, position, potential,
force, virial)` follow src/force/sus2_sh.cuh:12-25: outputs are accumulated (+=) into
caller-owned device arrays in GPUMD's SoA layout (position/force [3, N], virial [9, N] rows
xx yy zz xy xz yz yx zx zy).  Device memory comes from torch tensors or hipmem.DeviceArray
(plumbing only); all arithmetic happens in the HIP kernels of libsus2sh.so.
"""
import ctypes

import numpy as np

from ._capi import Box, ModelInfo, check, lib
from .structures import make_box


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _is_torch(t):
    return type(t).__module__.startswith("torch")


def _check_arrays(pairs):
    """Device arrays are torch cuda tensors or hipmem.DeviceArray objects."""
    for t, dt in pairs:
        if _is_torch(t):
            if str(t.dtype) != "torch." + dt or not t.is_cuda or not t.is_contiguous():
                raise ValueError("arguments must be contiguous device tensors of dtype " + dt)
        elif str(getattr(t, "dtype", "")) != dt:
            raise ValueError("arguments must be device arrays of dtype " + dt)


def _stream(t, stream):
    if stream is not None:
        return ctypes.c_void_p(stream)
    if _is_torch(t):
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return ctypes.c_void_p(0)


def box_struct(cell, pbc=(1, 1, 1)):
    h, p = make_box(cell, pbc)
    b = Box()
    for k in range(18):
        b.h[k] = float(h[k])
    for k in range(3):
        b.pbc[k] = int(p[k])
    return b


class SUS2_SH:
    """Handle to one SUS2-SH model on the current HIP device."""

    def __init__(self, file_potential, num_atoms=0, options=()):
        opts = [o.encode() for o in options]
        arr = (ctypes.c_char_p * max(1, len(opts)))(*opts)
        handle = ctypes.c_void_p()
        code = lib().sus2sh_create(str(file_potential).encode(), int(num_atoms), len(opts), arr,
                                   ctypes.byref(handle))
        check(code, None)
        self._h = handle

    def close(self):
        if getattr(self, "_h", None):
            lib().sus2sh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def rc(self):
        return lib().sus2sh_cutoff(self._h)

    def info(self):
        mi = ModelInfo()
        check(lib().sus2sh_get_model_info(self._h, ctypes.byref(mi)), self._h)
        return mi.as_dict()

    def stage_times(self):
        buf = (ctypes.c_float * 8)()
        n = lib().sus2sh_get_stage_times(self._h, buf, 8)
        return list(buf)[:n]

    def tuning(self):
        """Autotuner graph-kernel times {(variant, tile): ms} (variant 0 rows, 1 balanced
        streams, 2 round-robin streams; variant 3 is the lane program, keyed (3, shape);
        variant 4 the block program, keyed (4, shape) with shapes 0, 1, 2 = 8, 16, 4 waves)."""
        buf = (ctypes.c_float * 32)()
        n = lib().sus2sh_get_tuning(self._h, buf, 32)
        out = {(k // 5, 1 << (k % 5)): buf[k] for k in range(min(n, 15)) if buf[k] >= 0}
        for k in range(15, n):
            if buf[k] >= 0:
                out[(3, k - 15) if k < 27 else (4, k - 27)] = buf[k]
        return out

    def edge_shapes(self):
        """Edge-kernel launch shapes of the last compute and the autotuner's times:
        ({"basic": (threads, lanes per atom), "force": (...), "basic_mfma": 0 | 1 | 2},
         {("basic" | "force", threads, lanes per atom): ms, ("basic_mfma_fp32",): ms})."""
        shape = (ctypes.c_int * 5)()
        buf = (ctypes.c_float * 25)()
        n = lib().sus2sh_get_edge_tuning(self._h, shape, buf, 25)
        cand = [(t, npl) for t in (256, 128, 64) for npl in (1, 2, 4, 8)]
        times = {}
        for k in range(min(n, 24)):
            if buf[k] >= 0:
                times[("basic" if k < 12 else "force",) + cand[k % 12]] = buf[k]
        if n > 24 and buf[24] >= 0:
            times[("basic_mfma_fp32",)] = buf[24]
        return ({"basic": (shape[0], shape[1]), "force": (shape[2], shape[3]),
                 "basic_mfma": shape[4]}, times)

    def mean_neighbors(self):
        return lib().sus2sh_mean_neighbors(self._h)

    def check(self, stream=None):
        """sus2sh_check: synchronise and raise Sus2shError (status 5 = neighbour capacity
        exceeded, 2 = bad type) for what deferred-check evaluations latched since the last
        check (option sus2_sh_deferred_check=1)."""
        import torch
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(lib().sus2sh_check(self._h, ctypes.c_void_p(s)), self._h)

    def get_params(self):
        n = self.info()["param_count"]
        out = np.zeros(n)
        check(lib().sus2sh_get_params(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      n), self._h)
        return out

    def set_params(self, theta):
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        check(lib().sus2sh_set_params(self._h, theta.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      len(theta)), self._h)

    def compute(self, box, type, position, potential, force, virial, stream=None,
                param_grad=None):
        """Accumulate per-atom energy, forces and virials.

        box: Box struct (see box_struct), type: int32 [N] cuda tensor, position: float64
        [3, N] (contiguous SoA), potential: float64 [N], force: float64 [3, N], virial:
        float64 [9, N]; param_grad (optional): float64 [param_count] += dE/dtheta.
        """
        n = int(type.numel())
        _check_arrays(((type, "int32"), (position, "float64"), (potential, "float64"),
                       (force, "float64"), (virial, "float64")))
        if position.numel() != 3 * n or potential.numel() != n or force.numel() != 3 * n \
                or virial.numel() != 9 * n:
            raise ValueError("compute: inconsistent array sizes")
        s = _stream(type, stream)
        if param_grad is None:
            code = lib().sus2sh_compute(self._h, ctypes.byref(box), n, _ptr(type), _ptr(position),
                                        _ptr(potential), _ptr(force), _ptr(virial), s)
        else:
            code = lib().sus2sh_compute_param_grad(self._h, ctypes.byref(box), n, _ptr(type),
                                                   _ptr(position), _ptr(potential), _ptr(force),
                                                   _ptr(virial), _ptr(param_grad), s)
        check(code, self._h)


def box_array(cells, pbc=(1, 1, 1)):
    """ctypes array of Box structs for compute_batch (built once, reusable across calls)."""
    boxes = (Box * len(cells))()
    for k, c in enumerate(cells):
        boxes[k] = box_struct(c, pbc)
    return boxes


def compute_batch(pot, cells, atom_offsets, type, position, potential, force, virial,
                  pbc=(1, 1, 1), param_grad=None, stream=None):
    """Batched evaluation (sus2sh_compute_batch) of independent periodic structures whose atoms
    are concatenated in the SoA arrays; cells: list of 3x3 (rows = cell vectors), or the
    result of box_array(cells, pbc) (pbc is then ignored)."""
    if isinstance(cells, ctypes.Array):
        boxes, ns = cells, len(cells)
    else:
        ns = len(cells)
        boxes = box_array(cells, pbc)
    offs = np.ascontiguousarray(atom_offsets, dtype=np.int32)
    assert len(offs) == ns + 1
    s = _stream(type, stream)
    code = lib().sus2sh_compute_batch(pot._h, ns, boxes,
                                      offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                      _ptr(type), _ptr(position), _ptr(potential), _ptr(force),
                                      _ptr(virial),
                                      _ptr(param_grad) if param_grad is not None else None, s)
    check(code, pot._h)


def md_nve(pot, box, type, position, velocity, masses, dt_fs, nsteps, stream=None):
    """Velocity-Verlet NVE on the GPU (sus2sh_md_nve).  position/velocity: float64 [3, N]
    cuda tensors advanced in place; masses: per-species amu.  Returns (energy trace
    [nsteps+1, 2] of (potential, kinetic) in eV, potential[N], force[3,N], virial[9,N])."""
    n = int(type.numel())
    m = np.ascontiguousarray(masses, dtype=np.float64)
    if _is_torch(position):
        import torch
        e = torch.zeros(n, dtype=torch.float64, device=position.device)
        f = torch.zeros(3, n, dtype=torch.float64, device=position.device)
        v = torch.zeros(9, n, dtype=torch.float64, device=position.device)
    else:
        from .hipmem import DeviceArray
        e = DeviceArray.zeros((n,), np.float64)
        f = DeviceArray.zeros((3, n), np.float64)
        v = DeviceArray.zeros((9, n), np.float64)
    trace = np.zeros((nsteps + 1, 2))
    s = _stream(type, stream)
    code = lib().sus2sh_md_nve(pot._h, ctypes.byref(box), n, _ptr(type),
                               m.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _ptr(position),
                               _ptr(velocity), float(dt_fs), int(nsteps), _ptr(e), _ptr(f), _ptr(v),
                               trace.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), s)
    check(code, pot._h)
    return trace, e, f, v


def md_npt_mttk(pot, box, type, position, velocity, masses, dt_fs, nsteps, t_start, t_stop,
                p_start, p_stop, tperiod=100.0, pperiod=1000.0, couple="aniso", stream=None):
    """NPT with the MTK barostat and Nose-Hoover chains on the GPU (sus2sh_md_npt_mttk), the
    `ensemble npt_mttk temp T1 T2 aniso p1 p2 tperiod pperiod` step of the reference's
    benchmark protocol (docs/gpumd-sus2-sh-interface-plan.md:264-271).  box (a Box struct) and
    position/velocity (float64 [3, N] cuda tensors) are advanced in place; temperatures in K,
    pressures in GPa, periods in time steps; couple "iso" or "aniso".  Returns (trace
    [nsteps+1, 8] of (potential, kinetic, conserved extended energy, volume, temperature,
    pressure, cube root of the volume, target temperature), potential[N], force[3,N],
    virial[9,N])."""
    couples = {"iso": 0, "aniso": 1}
    if couple not in couples:
        raise ValueError(f"couple must be 'iso' or 'aniso', got {couple!r}")
    n = int(type.numel())
    m = np.ascontiguousarray(masses, dtype=np.float64)
    if _is_torch(position):
        import torch
        e = torch.zeros(n, dtype=torch.float64, device=position.device)
        f = torch.zeros(3, n, dtype=torch.float64, device=position.device)
        v = torch.zeros(9, n, dtype=torch.float64, device=position.device)
    else:
        from .hipmem import DeviceArray
        e = DeviceArray.zeros((n,), np.float64)
        f = DeviceArray.zeros((3, n), np.float64)
        v = DeviceArray.zeros((9, n), np.float64)
    trace = np.zeros((nsteps + 1, 8))
    s = _stream(type, stream)
    code = lib().sus2sh_md_npt_mttk(
        pot._h, ctypes.byref(box), n, _ptr(type),
        m.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _ptr(position), _ptr(velocity),
        float(dt_fs), int(nsteps), float(t_start), float(t_stop), float(p_start), float(p_stop),
        float(tperiod), float(pperiod), couples[couple], _ptr(e), _ptr(f), _ptr(v),
        trace.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), s)
    check(code, pot._h)
    return trace, e, f, v


def maxwell_velocities(masses_per_atom, temperature, seed):
    """Maxwell-Boltzmann velocities (A/fs) at `temperature` K with zero total momentum."""
    kB = 8.617333262e-5  # eV/K
    m = np.asarray(masses_per_atom, dtype=np.float64)
    rng = np.random.default_rng(seed)
    sd = np.sqrt(kB * temperature / m * 9.648533212331e-3)  # sqrt(kT/m) in A/fs
    v = rng.normal(size=(3, len(m))) * sd[None, :]
    p = (v * m[None, :]).sum(axis=1) / m.sum()
    v -= p[:, None]
    return v


def evaluate(model_path, cell, types, positions, options=(), pbc=(1, 1, 1)):
    """One evaluation from host arrays; returns numpy (energy[N], force[3,N], virial[9,N]).
    Device buffers come from hipmem (no torch needed)."""
    from .hipmem import DeviceArray, synchronize
    n = len(types)
    pot = SUS2_SH(model_path, n, options)
    t = DeviceArray.from_numpy(np.asarray(types, dtype=np.int32))
    x = DeviceArray.from_numpy(np.ascontiguousarray(positions, dtype=np.float64))
    e = DeviceArray.zeros((n,), np.float64)
    f = DeviceArray.zeros((3, n), np.float64)
    v = DeviceArray.zeros((9, n), np.float64)
    pot.compute(box_struct(cell, pbc), t, x, e, f, v)
    synchronize()
    out = (e.numpy(), f.numpy(), v.numpy())
    pot.close()
    return out
