"""Standard SUS2-SH graphs and seeded synthetic `.mtp` models.

The graph itself comes from the engine's C++ generator (C ABI sus2sh_standard_graph), which
restates SHStandardGraphBuilderHost (src/force/sus2_sh.cu:1227-1706).  This module only
draws coefficients and writes the text format read by the loader
(src/force/sus2_sh.cu:2105-2482): braced arrays, `alpha_index_basic = {{k, l, m}, ...}`,
`sh_products = {{left, right, target, coeff}, ...}` and a `radial_coeffs` section of
M_r rows with rb coefficients followed by S type coefficients.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from ._capi import check, lib


@dataclass
class StandardGraph:
    l_max: int
    k_max: int
    body_order: int
    body_l_max: tuple
    node_count: int
    basic_klm: np.ndarray      # [B, 3]
    products: np.ndarray       # [T, 3] int (left, right, target)
    coeffs: np.ndarray         # [T] float64
    scalars: np.ndarray        # [Q]
    scalar_body: np.ndarray    # [Q]
    max_layer: int

    @property
    def basic_count(self):
        return len(self.basic_klm)

    @property
    def product_count(self):
        return len(self.coeffs)

    @property
    def scalar_count(self):
        return len(self.scalars)


def standard_graph(l_max, k_max, body_order, body_l_max=None):
    """Build the standard real-CG graph with the engine's generator."""
    L = lib()
    handle = ctypes.c_void_p()
    if body_l_max is None:
        arr, n = None, 0
    else:
        body_l_max = tuple(int(v) for v in body_l_max)
        arr = (ctypes.c_int * len(body_l_max))(*body_l_max)
        n = len(body_l_max)
    check(L.sus2sh_standard_graph(l_max, k_max, body_order, arr, n, ctypes.byref(handle)))
    try:
        dims = (ctypes.c_int * 5)()
        check(L.sus2sh_graph_dims(handle, dims))
        A, B, T, Q, layers = list(dims)
        basic = np.zeros(3 * B, dtype=np.int32)
        prod = np.zeros(3 * max(T, 1), dtype=np.int32)
        coef = np.zeros(max(T, 1), dtype=np.float64)
        sc = np.zeros(max(Q, 1), dtype=np.int32)
        sb = np.zeros(max(Q, 1), dtype=np.int32)
        P = ctypes.POINTER
        check(L.sus2sh_graph_export(
            handle, basic.ctypes.data_as(P(ctypes.c_int)), prod.ctypes.data_as(P(ctypes.c_int)),
            coef.ctypes.data_as(P(ctypes.c_double)), sc.ctypes.data_as(P(ctypes.c_int)),
            sb.ctypes.data_as(P(ctypes.c_int))))
    finally:
        L.sus2sh_graph_free(handle)
    if body_l_max is None:
        body_l_max = tuple([l_max] * (5 if body_order >= 6 else 4))
    return StandardGraph(l_max, k_max, body_order, tuple(body_l_max), A, basic.reshape(B, 3),
                         prod[:3 * T].reshape(T, 3), coef[:T], sc[:Q], sb[:Q], layers)


def subgraph(g: StandardGraph, keep):
    """Explicit product graph that computes only the scalars selected by the boolean mask
    `keep` [Q]: the products those scalars depend on, with the non-basic moments renumbered
    in their original (topological) order.  The result is a valid non-standard graph of the
    kind the loader accepts through its explicit-graph path (src/force/sus2_sh.cu:1872-2103),
    e.g. a pruned (q-total) training graph."""
    keep = np.asarray(keep, dtype=bool)
    B = g.basic_count
    needed = np.zeros(g.node_count, dtype=bool)
    needed[g.scalars[keep]] = True
    used = np.zeros(g.product_count, dtype=bool)
    for t in range(g.product_count - 1, -1, -1):
        a, b, c = g.products[t]
        if needed[c]:
            used[t] = True
            needed[a] = needed[b] = True
    needed[:B] = True
    new_id = np.cumsum(needed) - 1
    prods = new_id[g.products[used]].astype(np.int32)
    layer = np.zeros(g.node_count, dtype=np.int64)
    for a, b, c in g.products[used]:
        layer[c] = max(layer[c], max(layer[a], layer[b]) + 1)
    return StandardGraph(g.l_max, g.k_max, g.body_order, g.body_l_max, int(needed.sum()),
                         g.basic_klm.copy(), prods, g.coeffs[used].copy(),
                         new_id[g.scalars[keep]].astype(np.int32), g.scalar_body[keep].copy(),
                         int(layer.max()))


@dataclass
class ModelSpec:
    """Everything needed to write a SUS2-SH `.mtp` file."""
    graph: StandardGraph
    species_count: int
    radial_basis_type: str
    rb_size: int
    max_dist: float
    scaling: float
    shift_coeffs: np.ndarray
    species_coeffs: np.ndarray
    scal_coeffs: np.ndarray          # [2 S^2 Mr]
    moment_coeffs: np.ndarray        # [Q]
    radial_coeffs: np.ndarray        # [Mr, rb]
    type_coeffs: np.ndarray          # [S]
    min_dist: float = 1.0
    name: str = "sus2sh_synthetic"
    extra: dict = field(default_factory=dict)
    # keys written after every other field, e.g. the two-layer gate block (its
    # two_layer_gate_radial_coeffs key must come after the radial_coeffs section, which the
    # loader finds by the first occurrence of "radial_coeffs")
    tail: dict = field(default_factory=dict)

    @property
    def radial_funcs_count(self):
        return self.graph.k_max * (self.graph.l_max + 1)


def synthetic_model(l_max, k_max, body_order=5, body_l_max=None, species_count=2, rb_size=8,
                    max_dist=5.0, radial_basis_type="RBChebyshev_sss", seed=0, coeff_scale=0.1,
                    scaling=1.0, type_coeff_mode="one_plus", scal_mode="uniform",
                    pair_core=None):
    """Seeded synthetic model on the standard graph.

    Coefficients: radial_coeffs, shift_coeffs, species_coeffs and moment_coeffs are drawn
    N(0, coeff_scale).  Type coefficients (trailing numbers of the mu=0 radial row) are
    1 + N(0, coeff_scale) with type_coeff_mode="one_plus" (pure N(0, coeff_scale) makes every
    pair scale ~1e-2) or N(0, coeff_scale) with "normal".  scal_coeffs: with
    scal_mode="uniform" every (mu, pair) uses a = 1, s = max_dist / 2; with "random"
    a ~ U[0.7, 1.3] and s ~ max_dist/2 + N(0, 0.3), independently per (mu, pair).
    pair_core (eV, optional): adds a repulsive pair core for molecular dynamics -- the
    l = 0 radial rows get c_0 += 1 (so R ~ (r - rc)^2, decreasing up to rc), the species
    coefficients get +1 and every body-2 scalar weight is set to +pair_core.  Purely random
    coefficients have no repulsion and heat up without bound under NVE.
    """
    g = standard_graph(l_max, k_max, body_order, body_l_max)
    rng = np.random.default_rng(seed)
    S = species_count
    Mr = k_max * (l_max + 1)
    radial = rng.normal(0.0, coeff_scale, size=(Mr, rb_size))
    if type_coeff_mode == "one_plus":
        tcoef = 1.0 + rng.normal(0.0, coeff_scale, size=S)
    else:
        tcoef = rng.normal(0.0, coeff_scale, size=S)
    shift = rng.normal(0.0, coeff_scale, size=S)
    species = rng.normal(0.0, coeff_scale, size=S)
    moment = rng.normal(0.0, coeff_scale, size=g.scalar_count)
    scal = np.zeros((Mr, 2, S * S))
    if scal_mode == "uniform":
        scal[:, 0, :] = 1.0
        scal[:, 1, :] = max_dist / 2.0
    else:
        scal[:, 0, :] = rng.uniform(0.7, 1.3, size=(Mr, S * S))
        scal[:, 1, :] = max_dist / 2.0 + rng.normal(0.0, 0.3, size=(Mr, S * S))
    if pair_core is not None:
        for mu in range(Mr):
            if mu % (l_max + 1) == 0:  # mu = k (L+1) + l with l = 0
                radial[mu, 0] += 1.0
        species = species + 1.0
        moment = np.where(g.scalar_body == 2, float(pair_core), moment)
    return ModelSpec(g, S, radial_basis_type, rb_size, max_dist, scaling, shift, species,
                     scal.reshape(-1), moment, radial, tcoef,
                     name=f"l{l_max}k{k_max}_b{body_order}_seed{seed}")


def _fmt(x):
    return repr(float(x))


def _braced(values, fmt=_fmt):
    return "{" + ", ".join(fmt(v) for v in values) + "}"


def write_mtp(spec: ModelSpec, path):
    """Write `spec` in the SUS2-SH `.mtp` text format (keys found at first occurrence)."""
    g = spec.graph
    lines = [
        "MTP",
        "version = 1.1.0",
        f"potential_name = {spec.name}",
        "potential_tag = SUS2-SH",
        f"scaling = {_fmt(spec.scaling)}",
        "scaling_map = LK",
        f"species_count = {spec.species_count}",
        f"sh_l_max = {g.l_max}",
        f"sh_k_max = {g.k_max}",
        f"sh_body_order = {g.body_order}",
        "sh_body_l_max = " + _braced(g.body_l_max, str),
        "sh_parity = even",
        f"radial_basis_type = {spec.radial_basis_type}",
        f"\tmin_dist = {_fmt(spec.min_dist)}",
        f"\tmax_dist = {_fmt(spec.max_dist)}",
    ]
    # optional header keys, e.g. the ZBL pair term: zbl_enabled, zbl_inner, zbl_outer,
    # zbl_typewise_cutoff_factor, zbl_atomic_numbers
    for key, val in spec.extra.items():
        if isinstance(val, bool):
            val = "true" if val else "false"
        elif isinstance(val, (list, tuple, np.ndarray)):
            val = _braced(list(val), str)
        elif isinstance(val, float):
            val = _fmt(val)
        lines.append(f"{key} = {val}")
    lines += [
        f"\tradial_basis_size = {spec.rb_size}",
        f"\tradial_funcs_count = {spec.radial_funcs_count}",
        "\tscal_coeffs = " + _braced(spec.scal_coeffs),
        "\tradial_coeffs",
    ]
    for mu in range(spec.radial_funcs_count):
        row = list(spec.radial_coeffs[mu]) + list(spec.type_coeffs)
        lines.append("\t\t" + _braced(row))
    lines += [
        f"alpha_moments_count = {g.node_count}",
        f"alpha_index_basic_count = {g.basic_count}",
        "alpha_index_basic = {" + ", ".join(
            "{%d, %d, %d}" % (k, l, m) for k, l, m in g.basic_klm) + "}",
        f"sh_product_count = {g.product_count}",
        "sh_products = {" + ", ".join(
            "{%d, %d, %d, %s}" % (a, b, t, _fmt(c))
            for (a, b, t), c in zip(g.products, g.coeffs)) + "}",
        f"alpha_scalar_moments = {g.scalar_count}",
        "alpha_moment_mapping = " + _braced(g.scalars, str),
        f"sh_scalar_info_count = {g.scalar_count}",
        "sh_scalar_info = {" + ", ".join("{%d}" % b for b in g.scalar_body) + "}",
        "shift_coeffs = " + _braced(spec.shift_coeffs),
        "species_coeffs = " + _braced(spec.species_coeffs),
        "moment_coeffs = " + _braced(spec.moment_coeffs),
    ]
    for key, val in spec.tail.items():
        if isinstance(val, bool):
            val = "true" if val else "false"
        elif isinstance(val, (list, tuple, np.ndarray)):
            val = _braced(list(val), _fmt if np.asarray(val).dtype.kind == "f" else str)
        elif isinstance(val, float):
            val = _fmt(val)
        lines.append(f"{key} = {val}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def gate_model(spec: ModelSpec, mode="mu-scalar-full", seed=0, amplitude=0.8, coeff_scale=0.1,
               scalars_per_body=None):
    """Copy of `spec` with a two-layer mu gate (docs/mu-body-order-gate.md:25-63): seeded gate
    radial coefficients N(0, coeff_scale), additive coefficients 1 + N(0, 0.2) per species,
    gate scalars = the model's scalars of body orders 2 .. k_max + 1 (at most
    `scalars_per_body` of each), weights N(0, coeff_scale) ([M_r][Qg] for mu-scalar-full,
    [Qg] plus body-mix weights N(0, 1) [M_r][k_max] for mu-body-linear-combo)."""
    import copy
    g = spec.graph
    rng = np.random.default_rng(seed)
    K, Mr, S = g.k_max, spec.radial_funcs_count, spec.species_count
    picks = []
    for body in range(2, K + 2):
        idx = np.flatnonzero(g.scalar_body == body)
        if scalars_per_body is not None:
            idx = idx[:scalars_per_body]
        picks.extend(int(i) for i in idx)
    Qg = len(picks)
    out = copy.deepcopy(spec)
    tail = {
        "two_layer_gate_enabled": True,
        "two_layer_gate_mode": mode,
        "two_layer_gate_site_mode": "neighbor",
        "two_layer_gate_radial_mode": "shared-radial",
        "two_layer_gate_tanh_amplitude": float(amplitude),
        "two_layer_gate_body_order_max": K + 1,
        "two_layer_gate_radial_coeff_count": Mr * spec.rb_size,
        "two_layer_gate_radial_coeffs": rng.normal(0.0, coeff_scale, size=Mr * spec.rb_size),
        "two_layer_gate_additive_coeff_count": S,
        "two_layer_gate_additive_coeffs": 1.0 + rng.normal(0.0, 0.2, size=S),
    }
    if mode == "mu-body-linear-combo":
        tail["two_layer_gate_weight_count"] = Qg
        tail["two_layer_gate_scalar_indices"] = np.array(picks, dtype=np.int64)
        tail["two_layer_gate_weights"] = rng.normal(0.0, coeff_scale, size=Qg)
        tail["two_layer_gate_body_mix_weight_count"] = Mr * K
        tail["two_layer_gate_body_mix_weights"] = rng.normal(0.0, 1.0, size=Mr * K)
    else:
        tail["two_layer_gate_weight_count"] = Mr * Qg
        tail["two_layer_gate_scalar_indices"] = np.array(picks, dtype=np.int64)
        tail["two_layer_gate_weights"] = rng.normal(0.0, coeff_scale, size=Mr * Qg)
    out.tail = tail
    out.name = spec.name + "_gate"
    return out


def param_vector(spec: ModelSpec):
    """Flat parameter vector in the engine's dE/dtheta order:
    [shift S][species S][moment Q][radial Mr*rb][type S][scal 2 S^2 Mr]."""
    return np.concatenate([spec.shift_coeffs, spec.species_coeffs, spec.moment_coeffs,
                           spec.radial_coeffs.reshape(-1), spec.type_coeffs, spec.scal_coeffs])


def with_params(spec: ModelSpec, theta):
    """Copy of spec with parameters replaced from a flat vector (param_vector order)."""
    import copy
    S = spec.species_count
    Q = spec.graph.scalar_count
    Mr, rb = spec.radial_coeffs.shape
    out = copy.deepcopy(spec)
    o = 0
    out.shift_coeffs = np.array(theta[o:o + S]); o += S
    out.species_coeffs = np.array(theta[o:o + S]); o += S
    out.moment_coeffs = np.array(theta[o:o + Q]); o += Q
    out.radial_coeffs = np.array(theta[o:o + Mr * rb]).reshape(Mr, rb); o += Mr * rb
    out.type_coeffs = np.array(theta[o:o + S]); o += S
    out.scal_coeffs = np.array(theta[o:o + 2 * S * S * Mr]); o += 2 * S * S * Mr
    assert o == len(theta)
    return out
