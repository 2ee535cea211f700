"""ctypes binding of the C ABI declared in include/sus2sh.h.

The shared library is built in-tree (gpumd-sus2-sh_amd/sus2sh/libsus2sh.so).  Loading
fails loudly when it is missing: there is no CPU fallback for the engine.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsus2sh.so")

SUS2SH_OK = 0
SUS2SH_ERR_INPUT = 1
SUS2SH_ERR_ARGUMENT = 2
SUS2SH_ERR_DEVICE = 3
SUS2SH_ERR_UNSUPPORTED = 4
SUS2SH_ERR_CAPACITY = 5

# Every symbol declared in include/sus2sh.h.
EXPORTED_SYMBOLS = (
    "sus2sh_probe",
    "sus2sh_create",
    "sus2sh_destroy",
    "sus2sh_compute",
    "sus2sh_compute_noalloc",
    "sus2sh_compute_param_grad",
    "sus2sh_compute_batch",
    "sus2sh_md_nve",
    "sus2sh_md_npt_mttk",
    "sus2sh_check",
    "sus2sh_set_params",
    "sus2sh_get_params",
    "sus2sh_last_error",
    "sus2sh_cutoff",
    "sus2sh_get_model_info",
    "sus2sh_get_stage_times",
    "sus2sh_get_tuning",
    "sus2sh_get_edge_tuning",
    "sus2sh_mean_neighbors",
    "sus2sh_clebsch_gordan",
    "sus2sh_real_cg",
    "sus2sh_standard_graph",
    "sus2sh_graph_dims",
    "sus2sh_graph_export",
    "sus2sh_graph_free",
    "sus2sh_model_check",
    "sus2sh_lane_check",
    "sus2sh_block_check",
    "sus2sh_zbl_tables",
)


class Sus2shError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[sus2sh status {code}] {message}")
        self.code = code
        self.message = message


class Box(ctypes.Structure):
    _fields_ = [("h", ctypes.c_double * 18), ("pbc", ctypes.c_int * 3)]


class ModelInfo(ctypes.Structure):
    _fields_ = [
        ("species_count", ctypes.c_int),
        ("l_max", ctypes.c_int),
        ("k_max", ctypes.c_int),
        ("body_order", ctypes.c_int),
        ("radial_basis_size", ctypes.c_int),
        ("radial_funcs_count", ctypes.c_int),
        ("moments_count", ctypes.c_int),
        ("basic_count", ctypes.c_int),
        ("scalar_count", ctypes.c_int),
        ("product_count", ctypes.c_int),
        ("layer_count", ctypes.c_int),
        ("row_count", ctypes.c_int),
        ("back_row_count", ctypes.c_int),
        ("back_term_count", ctypes.c_int),
        ("standard_graph", ctypes.c_int),
        ("use_float", ctypes.c_int),
        ("param_count", ctypes.c_int),
        ("stored_slots", ctypes.c_int),
        ("graph_tile", ctypes.c_int),
        ("graph_variant", ctypes.c_int),
        ("neighbor_capacity", ctypes.c_int),
        ("max_dist", ctypes.c_double),
        ("lane_program", ctypes.c_int),
        ("lane_slots", ctypes.c_int),
        ("lane_instructions", ctypes.c_int),
        ("lane_waves", ctypes.c_int),
        ("lane_nslot", ctypes.c_int),
        ("lane_atoms", ctypes.c_int),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None


def _load_torch_runtime_first():
    """torch bundles its own HIP runtime with the same soname (libamdhip64.so.7) as the system
    runtime libsus2sh.so links; whichever is loaded first serves both.  Importing torch first
    keeps torch working in processes that use both."""
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def lib():
    """Load libsus2sh.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    _load_torch_runtime_first()
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C gpumd-sus2-sh_amd` "
            "(or __graft_entry__.build()); the SUS2-SH engine has no fallback path")
    L = ctypes.CDLL(LIB_PATH)
    c_int, c_double, c_char_p, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_char_p, ctypes.c_void_p
    P = ctypes.POINTER
    L.sus2sh_probe.argtypes = [c_char_p]
    L.sus2sh_probe.restype = c_int
    L.sus2sh_create.argtypes = [c_char_p, c_int, c_int, P(c_char_p), P(c_void_p)]
    L.sus2sh_create.restype = c_int
    L.sus2sh_destroy.argtypes = [c_void_p]
    L.sus2sh_destroy.restype = None
    L.sus2sh_compute.argtypes = [c_void_p, P(Box), c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p]
    L.sus2sh_compute.restype = c_int
    L.sus2sh_compute_noalloc.argtypes = [c_void_p, P(Box), c_int, c_void_p, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p]
    L.sus2sh_compute_noalloc.restype = c_int
    L.sus2sh_compute_param_grad.argtypes = [c_void_p, P(Box), c_int, c_void_p, c_void_p, c_void_p,
                                            c_void_p, c_void_p, c_void_p, c_void_p]
    L.sus2sh_compute_param_grad.restype = c_int
    L.sus2sh_compute_batch.argtypes = [c_void_p, c_int, P(Box), P(c_int), c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
    L.sus2sh_compute_batch.restype = c_int
    L.sus2sh_md_nve.argtypes = [c_void_p, P(Box), c_int, c_void_p, P(c_double), c_void_p,
                                c_void_p, c_double, c_int, c_void_p, c_void_p, c_void_p,
                                P(c_double), c_void_p]
    L.sus2sh_md_nve.restype = c_int
    L.sus2sh_md_npt_mttk.argtypes = [c_void_p, P(Box), c_int, c_void_p, P(c_double), c_void_p,
                                     c_void_p, c_double, c_int, c_double, c_double, c_double,
                                     c_double, c_double, c_double, c_int, c_void_p, c_void_p,
                                     c_void_p, P(c_double), c_void_p]
    L.sus2sh_md_npt_mttk.restype = c_int
    L.sus2sh_check.argtypes = [c_void_p, c_void_p]
    L.sus2sh_check.restype = c_int
    L.sus2sh_set_params.argtypes = [c_void_p, P(c_double), c_int]
    L.sus2sh_set_params.restype = c_int
    L.sus2sh_get_params.argtypes = [c_void_p, P(c_double), c_int]
    L.sus2sh_get_params.restype = c_int
    L.sus2sh_last_error.argtypes = [c_void_p]
    L.sus2sh_last_error.restype = c_char_p
    L.sus2sh_cutoff.argtypes = [c_void_p]
    L.sus2sh_cutoff.restype = c_double
    L.sus2sh_get_model_info.argtypes = [c_void_p, P(ModelInfo)]
    L.sus2sh_get_model_info.restype = c_int
    L.sus2sh_get_stage_times.argtypes = [c_void_p, P(ctypes.c_float), c_int]
    L.sus2sh_get_stage_times.restype = c_int
    L.sus2sh_get_tuning.argtypes = [c_void_p, P(ctypes.c_float), c_int]
    L.sus2sh_get_tuning.restype = c_int
    L.sus2sh_get_edge_tuning.argtypes = [c_void_p, P(c_int), P(ctypes.c_float), c_int]
    L.sus2sh_get_edge_tuning.restype = c_int
    L.sus2sh_mean_neighbors.argtypes = [c_void_p]
    L.sus2sh_mean_neighbors.restype = c_double
    L.sus2sh_clebsch_gordan.argtypes = [c_int] * 6
    L.sus2sh_clebsch_gordan.restype = c_double
    L.sus2sh_real_cg.argtypes = [c_int] * 6
    L.sus2sh_real_cg.restype = c_double
    L.sus2sh_standard_graph.argtypes = [c_int, c_int, c_int, P(c_int), c_int, P(c_void_p)]
    L.sus2sh_standard_graph.restype = c_int
    L.sus2sh_graph_dims.argtypes = [c_void_p, P(c_int)]
    L.sus2sh_graph_dims.restype = c_int
    L.sus2sh_graph_export.argtypes = [c_void_p, P(c_int), P(c_int), P(c_double), P(c_int), P(c_int)]
    L.sus2sh_graph_export.restype = c_int
    L.sus2sh_graph_free.argtypes = [c_void_p]
    L.sus2sh_graph_free.restype = None
    L.sus2sh_model_check.argtypes = [c_char_p, P(c_int)]
    L.sus2sh_model_check.restype = c_int
    L.sus2sh_lane_check.argtypes = [c_char_p, c_int, c_int, c_int, c_int, c_int, c_int,
                                    P(c_double), P(c_double), P(c_double), P(c_int)]
    L.sus2sh_lane_check.restype = c_int
    L.sus2sh_block_check.argtypes = [c_char_p, c_int, c_int, c_int, c_int, c_int, P(c_double),
                                     P(c_double), P(c_double), P(c_int)]
    L.sus2sh_block_check.restype = c_int
    L.sus2sh_zbl_tables.argtypes = [c_char_p, c_int, P(c_char_p), P(c_int), P(c_double),
                                    P(c_double)]
    L.sus2sh_zbl_tables.restype = c_int
    _lib = L
    return L


def last_error(handle=None):
    msg = lib().sus2sh_last_error(handle)
    return msg.decode() if msg else ""


def check(code, handle=None):
    if code != SUS2SH_OK:
        raise Sus2shError(code, last_error(handle))
