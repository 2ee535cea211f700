"""MI355X-native SUS2-SH potential engine (host-side Python mirror of the C ABI).

The compute path is libsus2sh.so (HIP kernels for gfx950); this package only binds it.
"""
from ._capi import EXPORTED_SYMBOLS, LIB_PATH, Box, Sus2shError, lib, last_error  # noqa: F401
from .model_gen import (ModelSpec, StandardGraph, gate_model, param_vector, standard_graph,  # noqa: F401
                        subgraph, synthetic_model, with_params, write_mtp)
from .potential import (SUS2_SH, box_array, box_struct, compute_batch, evaluate,  # noqa: F401
                        maxwell_velocities, md_npt_mttk, md_nve)
from . import structures  # noqa: F401
from . import gpumd_io  # noqa: F401


def probe(path):
    """1 when `path` is a SUS2-SH `.mtp` file (first token MTP, potential_tag = SUS2-SH)."""
    return bool(lib().sus2sh_probe(str(path).encode()))


def clebsch_gordan(j1, m1, j2, m2, j, m):
    return lib().sus2sh_clebsch_gordan(j1, m1, j2, m2, j, m)


def real_cg(l1, m1, l2, m2, L, M):
    return lib().sus2sh_real_cg(l1, m1, l2, m2, L, M)


def model_check(path):
    """Host-side load + validation; returns a dict of the parsed dimensions."""
    import ctypes
    from ._capi import check
    dims = (ctypes.c_int * 16)()
    check(lib().sus2sh_model_check(str(path).encode(), dims))
    keys = ("species_count", "l_max", "k_max", "body_order", "radial_basis_size",
            "radial_funcs_count", "moments_count", "basic_count", "scalar_count",
            "product_count", "layer_count", "row_count", "row_term_count", "back_row_count",
            "back_term_count", "standard_graph")
    return dict(zip(keys, list(dims)))


def lane_check(path, basics=None, max_slots=160, waves=1, nslot=0, real_bytes=0, atoms=1):
    """Host-only run of the lane program (the one-atom-per-lane graph kernel's instruction
    stream) for basic vectors basics [n, B], compiled for `waves` waves per tile and `nslot`
    accumulators per wave (0: 2 with 8 waves, else 4).  real_bytes 0 runs the compiled
    program, 4 / 8 its FP32 / FP64 device encoding for `atoms` atoms per lane.  Returns (info dict, e_prime [n],
    grad [n, B]); e_prime / grad are None when the model has no lane program."""
    import ctypes
    import numpy as np
    from ._capi import check
    P = ctypes.POINTER(ctypes.c_double)
    info = (ctypes.c_int * 6)()
    args = (str(path).encode(), int(max_slots), int(waves), int(nslot), int(real_bytes),
            int(atoms))
    if basics is None:
        check(lib().sus2sh_lane_check(*args, 0, None, None, None, info))
        e = g = None
    else:
        b = np.ascontiguousarray(np.atleast_2d(basics), dtype=np.float64)
        e = np.zeros(b.shape[0])
        g = np.zeros_like(b)
        check(lib().sus2sh_lane_check(*args, b.shape[0], b.ctypes.data_as(P),
                                      e.ctypes.data_as(P), g.ctypes.data_as(P), info))
    keys = ("ok", "slots", "instructions", "components", "batches", "largest_component")
    d = dict(zip(keys, list(info)))
    if not d["ok"]:
        e = g = None
    return d, e, g


def block_check(path, basics=None, max_slots=1024, waves=8, real_bytes=4, target_slots=0):
    """Host-only run of the block program (the block graph kernel's device words, graph
    variant 4) for basic vectors basics [n, B]; target_slots > 0 halves the large degree
    classes when that brings the program to that many per-lane slots.  Returns (info dict,
    e_prime [n], grad [n, B]); e_prime / grad are None when the model has no block program."""
    import ctypes
    import numpy as np
    from ._capi import check
    P = ctypes.POINTER(ctypes.c_double)
    info = (ctypes.c_int * 9)()
    args = (str(path).encode(), int(max_slots), int(target_slots), int(waves), int(real_bytes))
    if basics is None:
        check(lib().sus2sh_block_check(*args, 0, None, None, None, info))
        e = g = None
    else:
        b = np.ascontiguousarray(np.atleast_2d(basics), dtype=np.float64)
        e = np.zeros(b.shape[0])
        g = np.zeros_like(b)
        check(lib().sus2sh_block_check(*args, b.shape[0], b.ctypes.data_as(P),
                                       e.ctypes.data_as(P), g.ctypes.data_as(P), info))
    keys = ("ok", "slots", "batches", "tensors", "fwd_blocks", "fwd_terms", "adj_terms",
            "bwd_terms", "units")
    d = dict(zip(keys, list(info)))
    if not d["ok"]:
        e = g = None
    return d, e, g


def zbl_tables(path, symbols=()):
    """Host-only ZBL tables of a zbl_enabled model (sus2sh_zbl_tables): (atomic numbers [S],
    inner [S, S], outer [S, S]); `symbols` are the element symbols that follow the model file
    in run.in (used when the file has no zbl_atomic_numbers)."""
    import ctypes
    import numpy as np
    from ._capi import check
    dims = model_check(path)
    S = dims["species_count"]
    z = (ctypes.c_int * S)()
    a = np.zeros((S, S))
    b = np.zeros((S, S))
    arr = (ctypes.c_char_p * max(1, len(symbols)))(*[s.encode() for s in symbols])
    P = ctypes.POINTER(ctypes.c_double)
    check(lib().sus2sh_zbl_tables(str(path).encode(), len(symbols), arr, z, a.ctypes.data_as(P),
                                  b.ctypes.data_as(P)))
    return np.array(list(z)), a, b
