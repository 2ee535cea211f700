"""Python mirror of the GPUMD plugin interface, on top of the C ABI.

`SUS2_SH(file, num_atoms, options)` and `SUS2_SH.compute(box, type, position, potential,
force, virial)` follow src/force/sus2_sh.cuh:12-25: outputs are accumulated (+=) into
caller-owned device arrays in GPUMD's SoA layout (position/force [3, N], virial [9, N] rows
xx yy zz xy xz yz yx zx zy).  Device memory is provided by torch tensors (plumbing only);
all arithmetic happens in the HIP kernels of libsus2sh.so.
"""
import ctypes

import numpy as np

from ._capi import Box, ModelInfo, check, lib
from .structures import make_box


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


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

    def mean_neighbors(self):
        return lib().sus2sh_mean_neighbors(self._h)

    def compute(self, box, type, position, potential, force, virial, stream=None,
                param_grad=None):
        """Accumulate per-atom energy, forces and virials.

        box: Box struct (see box_struct), type: int32 [N] cuda tensor, position: float64
        [3, N] (contiguous SoA), potential: float64 [N], force: float64 [3, N], virial:
        float64 [9, N]; param_grad (optional): float64 [param_count] += dE/dtheta.
        """
        import torch
        n = int(type.numel())
        for t, dt in ((type, torch.int32), (position, torch.float64), (potential, torch.float64),
                      (force, torch.float64), (virial, torch.float64)):
            if t.dtype != dt or not t.is_cuda or not t.is_contiguous():
                raise ValueError("compute: arguments must be contiguous cuda tensors of the "
                                 "documented dtypes")
        if position.numel() != 3 * n or potential.numel() != n or force.numel() != 3 * n \
                or virial.numel() != 9 * n:
            raise ValueError("compute: inconsistent array sizes")
        s = ctypes.c_void_p(stream if stream is not None else
                            torch.cuda.current_stream().cuda_stream)
        if param_grad is None:
            code = lib().sus2sh_compute(self._h, ctypes.byref(box), n, _ptr(type), _ptr(position),
                                        _ptr(potential), _ptr(force), _ptr(virial), s)
        else:
            code = lib().sus2sh_compute_param_grad(self._h, ctypes.byref(box), n, _ptr(type),
                                                   _ptr(position), _ptr(potential), _ptr(force),
                                                   _ptr(virial), _ptr(param_grad), s)
        check(code, self._h)


def md_nve(pot, box, type, position, velocity, masses, dt_fs, nsteps, stream=None):
    """Velocity-Verlet NVE on the GPU (sus2sh_md_nve).  position/velocity: float64 [3, N]
    cuda tensors advanced in place; masses: per-species amu.  Returns (energy trace
    [nsteps+1, 2] of (potential, kinetic) in eV, potential[N], force[3,N], virial[9,N])."""
    import torch
    n = int(type.numel())
    m = np.ascontiguousarray(masses, dtype=np.float64)
    e = torch.zeros(n, dtype=torch.float64, device=position.device)
    f = torch.zeros(3, n, dtype=torch.float64, device=position.device)
    v = torch.zeros(9, n, dtype=torch.float64, device=position.device)
    trace = np.zeros((nsteps + 1, 2))
    s = ctypes.c_void_p(stream if stream is not None else torch.cuda.current_stream().cuda_stream)
    code = lib().sus2sh_md_nve(pot._h, ctypes.byref(box), n, _ptr(type),
                               m.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _ptr(position),
                               _ptr(velocity), float(dt_fs), int(nsteps), _ptr(e), _ptr(f), _ptr(v),
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


def evaluate(model_path, cell, types, positions, options=(), pbc=(1, 1, 1), device="cuda"):
    """One evaluation from host arrays; returns numpy (energy[N], force[3,N], virial[9,N])."""
    import torch
    n = len(types)
    pot = SUS2_SH(model_path, n, options)
    t = torch.as_tensor(np.asarray(types, dtype=np.int32), device=device)
    x = torch.as_tensor(np.ascontiguousarray(positions, dtype=np.float64), device=device)
    e = torch.zeros(n, dtype=torch.float64, device=device)
    f = torch.zeros(3, n, dtype=torch.float64, device=device)
    v = torch.zeros(9, n, dtype=torch.float64, device=device)
    pot.compute(box_struct(cell, pbc), t, x, e, f, v)
    torch.cuda.synchronize()
    out = (e.cpu().numpy(), f.cpu().numpy(), v.cpu().numpy())
    pot.close()
    return out
