"""GPU parity: the HIP engine (through the C ABI) against the FP64 CPU oracle.

Tolerances (stated per precision):
  FP64 engine  : |dE| <= 1e-9 eV/atom, |dF| <= 1e-8 eV/A, |dW| <= 1e-8 eV (the whole path is
                 double; the north-star bound is 1e-6 eV/atom and 1e-5 eV/A)
  FP32 engine  : |dE| <= 2e-5 eV/atom + 2e-6 |E|, |dF| <= 1e-4 eV/A + 2e-5 |F|max,
                 |dW| <= 1e-4 eV + 2e-5 |W|max
"""
import numpy as np
import pytest

import sus2sh
from sus2sh import structures as st
from oracle.binding import Oracle

pytestmark = pytest.mark.gpu

TOL64 = dict(e=1e-9, f=1e-8, w=1e-8)


def _run(path, cell, types, pos, options=(), pbc=(1, 1, 1)):
    print(f"[case] model={path} N={len(types)} options={options}", flush=True)
    return sus2sh.evaluate(path, cell, types, pos, options=options, pbc=pbc)


def _compare(ref, got, fp32):
    e0, f0, v0 = ref
    e1, f1, v1 = got
    de = np.abs(e1 - e0).max()
    df = np.abs(f1 - f0).max()
    dv = np.abs(v1 - v0).max()
    print(f"  max|dE|={de:.3e} eV/atom  max|dF|={df:.3e} eV/A  max|dW|={dv:.3e}  "
          f"(|F|max={np.abs(f0).max():.3f}, |E|max={np.abs(e0).max():.3f})", flush=True)
    if fp32:
        assert de <= 2e-5 + 2e-6 * np.abs(e0).max()
        assert df <= 1e-4 + 2e-5 * np.abs(f0).max()
        assert dv <= 1e-4 + 2e-5 * np.abs(v0).max()
    else:
        assert de <= TOL64["e"]
        assert df <= TOL64["f"]
        assert dv <= TOL64["w"]


CASES = {
    "l2k2_fcc": dict(model=dict(l_max=2, k_max=2, body_order=5, rb_size=8, max_dist=5.0, seed=11),
                     struct=("fcc", 4, 3.80, 0.05)),
    "l3k3_fcc": dict(model=dict(l_max=3, k_max=3, body_order=5, rb_size=8, max_dist=5.0, seed=12,
                                scaling=0.5),
                     struct=("fcc", 5, 3.85, 0.08)),
    "l4k4_4422": dict(model=dict(l_max=4, k_max=4, body_order=5, body_l_max=(4, 4, 2, 2),
                                 rb_size=10, max_dist=5.5, seed=13, species_count=3,
                                 radial_basis_type="RBChebyshev_sss_rational", scaling=0.4),
                      struct=("fcc", 4, 3.85, 0.08)),
    "l3k2_nonuniform": dict(model=dict(l_max=3, k_max=2, body_order=4, rb_size=6, max_dist=4.6,
                                       seed=14, scal_mode="random"),
                            struct=("fcc", 4, 3.7, 0.1)),
    "l2k3_body6": dict(model=dict(l_max=2, k_max=3, body_order=6, rb_size=5, max_dist=4.2,
                                  seed=15, scaling=0.6),
                       struct=("fcc", 4, 3.6, 0.05)),
    # rb 14: the 16-coefficient uniform radial rows (rb <= 8 pads to 8, rb <= 12 to 12)
    "l2k2_rb14": dict(model=dict(l_max=2, k_max=2, body_order=5, rb_size=14, max_dist=5.0,
                                 seed=18, scaling=0.5),
                      struct=("fcc", 4, 3.85, 0.08)),
    "l5k2_5511": dict(model=dict(l_max=5, k_max=2, body_order=5, body_l_max=(5, 5, 1, 1),
                                 rb_size=10, max_dist=5.0, seed=16, scaling=0.5),
                      struct=("fcc", 4, 3.85, 0.08)),
    "l6k1_6611": dict(model=dict(l_max=6, k_max=1, body_order=5, body_l_max=(6, 6, 1, 1),
                                 rb_size=10, max_dist=5.0, seed=17, scaling=0.5),
                      struct=("fcc", 4, 3.85, 0.08)),
    # k_max > 6: the edge kernels run on k chunks (8 = 4 + 4, 13 = 5 + 4 + 4; radial
    # funcs 32 and 39 of the loader's 48)
    "l3k8_chunks": dict(model=dict(l_max=3, k_max=8, body_order=4, rb_size=8, max_dist=5.0,
                                   seed=19, scaling=0.4),
                        struct=("fcc", 4, 3.85, 0.08)),
    "l2k13_chunks": dict(model=dict(l_max=2, k_max=13, body_order=3, rb_size=6, max_dist=4.6,
                                    seed=20, scaling=0.4, scal_mode="random"),
                         struct=("fcc", 4, 3.7, 0.08)),
}


def _structure(spec, S, seed=7):
    kind, n, a, sigma = spec
    fr = [1.0 / S] * S
    return st.fcc_alloy(n, n, n, a, fr, sigma, seed, seed + 1)


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("fp32", [False, True])
def test_engine_matches_oracle(make_model, name, fp32):
    case = CASES[name]
    path = make_model(name, **case["model"])
    S = case["model"].get("species_count", 2)
    pos, types, cell = _structure(case["struct"], S)
    ref = Oracle(path).compute(cell, types, pos)
    opts = [f"T{s}" for s in range(S)] + ["sus2_float=%d" % (1 if fp32 else 0)]
    got = _run(path, cell, types, pos, options=opts)
    _compare(ref, got, fp32)


def _separated_positions(cell, n, dmin, rng):
    """Random positions in a periodic cell with every image pair at least dmin apart."""
    shifts = np.array([[i, j, k] for i in range(-3, 4) for j in range(-3, 4)
                       for k in range(-3, 4)], dtype=np.float64) @ cell
    pts = []
    while len(pts) < n:
        p = rng.uniform(0, 1, size=3) @ cell
        ok = True
        for q in pts + [p]:
            d = (q - p)[None, :] + shifts
            r = np.sqrt((d * d).sum(1))
            if q is p:
                r = r[r > 1e-12]
            if r.min() < dmin:
                ok = False
                break
        if ok:
            pts.append(p)
    return np.array(pts).T.copy()


@pytest.mark.parametrize("fp32", [False, True])
def test_small_triclinic_cells_multi_image(make_model, fp32):
    """Cells thinner than 2 rc (down to < rc) need every periodic image within rc
    (src/force/sus2_sh.cu:7668-7767)."""
    path = make_model("l3k3_small", l_max=3, k_max=3, body_order=5, rb_size=8, max_dist=5.0,
                      seed=21, scaling=0.5)
    orc = Oracle(path)
    rng = np.random.default_rng(5)
    for n in (8, 17, 40):
        cell = st.triclinic_cell(n, 0.065, rng)
        pos = _separated_positions(cell, n, 1.8, rng)
        types = rng.integers(0, 2, size=n).astype(np.int32)
        ref = orc.compute(cell, types, pos, brute=True)
        got = _run(path, cell, types, pos, options=["Cu", "Zr", "sus2_float=%d" % fp32])
        _compare(ref, got, fp32)


def test_open_boundaries(make_model):
    path = make_model("l2k2_open", l_max=2, k_max=2, body_order=4, rb_size=6, max_dist=5.0, seed=22)
    pos, types, cell = st.fcc_alloy(3, 3, 3, 3.8, [0.5, 0.5], 0.05, 3, 4)
    for pbc in ((0, 0, 0), (1, 0, 1)):
        ref = Oracle(path).compute(cell, types, pos, pbc=pbc, brute=True)
        got = _run(path, cell, types, pos, options=["A", "B", "sus2_float=0"], pbc=pbc)
        _compare(ref, got, False)


def test_accumulate_and_resize(make_model):
    """Outputs are accumulated (+=) and N may change between calls on one handle
    (src/force/force.cu:513-539; src/force/sus2_sh.cu:9628-9629)."""
    import torch
    path = make_model("l2k2_acc", l_max=2, k_max=2, body_order=5, rb_size=8, max_dist=5.0, seed=23)
    pot = sus2sh.SUS2_SH(path, 0, ["Cu", "Zr", "sus2_float=0"])
    orc = Oracle(path)
    for n in (3, 4, 3):
        pos, types, cell = st.fcc_alloy(n, n, n, 3.8, [0.5, 0.5], 0.05, n, n + 1)
        N = len(types)
        print(f"[case] accumulate N={N}", flush=True)
        ref = orc.compute(cell, types, pos)
        t = torch.as_tensor(types, device="cuda")
        x = torch.as_tensor(pos, device="cuda")
        e = torch.ones(N, dtype=torch.float64, device="cuda")
        f = torch.ones(3, N, dtype=torch.float64, device="cuda")
        v = torch.ones(9, N, dtype=torch.float64, device="cuda")
        box = sus2sh.box_struct(cell)
        pot.compute(box, t, x, e, f, v)
        pot.compute(box, t, x, e, f, v)
        torch.cuda.synchronize()
        assert np.abs(e.cpu().numpy() - (1 + 2 * ref[0])).max() < 1e-8
        assert np.abs(f.cpu().numpy() - (1 + 2 * ref[1])).max() < 1e-8
        assert np.abs(v.cpu().numpy() - (1 + 2 * ref[2])).max() < 1e-8
    # empty system is a no-op
    z = torch.zeros(0, dtype=torch.float64, device="cuda")
    pot.compute(sus2sh.box_struct(np.eye(3) * 20), torch.zeros(0, dtype=torch.int32, device="cuda"),
                z, z, z, z)
    pot.close()


def test_invalid_model_reports_error(tmp_path):
    bad = tmp_path / "bad.mtp"
    bad.write_text("MTP\nversion = 1.1.0\npotential_tag = SUS2-SH\n")
    with pytest.raises(sus2sh.Sus2shError) as err:
        sus2sh.SUS2_SH(str(bad), 10, ["Cu"])
    assert err.value.code == 1
    assert "Missing token" in err.value.message


PG_MODELS = {
    "l3k2": dict(l_max=3, k_max=2, body_order=5, rb_size=6, max_dist=4.8, seed=41,
                 scal_mode="random", species_count=3, scaling=0.5),
    # k chunks 4 + 4: the radial adjoints of each chunk land in its own block
    "l2k8": dict(l_max=2, k_max=8, body_order=3, rb_size=6, max_dist=4.8, seed=42,
                 scal_mode="random", species_count=3, scaling=0.4),
}


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("model", sorted(PG_MODELS))
def test_parameter_gradient_matches_oracle(make_model, fp32, model):
    """dE/dtheta over every model coefficient (no reference counterpart) against the FP64
    oracle's adjoint; the oracle's adjoint is itself pinned by finite differences on CPU."""
    import torch
    path = make_model(model + "_pg", **PG_MODELS[model])
    pos, types, cell = st.fcc_alloy(3, 3, 3, 3.7, [0.4, 0.4, 0.2], 0.08, 5, 6)
    ref = Oracle(path).compute(cell, types, pos, param_grad=True)
    pg_ref = ref[3]
    pot = sus2sh.SUS2_SH(path, len(types), ["A", "B", "C", "sus2_float=%d" % fp32])
    P = pot.info()["param_count"]
    assert P == len(pg_ref)
    N = len(types)
    print(f"[case] param grad {model} N={N} P={P} fp32={fp32}", flush=True)
    t = torch.as_tensor(types, device="cuda")
    x = torch.as_tensor(pos, device="cuda")
    e = torch.zeros(N, dtype=torch.float64, device="cuda")
    f = torch.zeros(3, N, dtype=torch.float64, device="cuda")
    v = torch.zeros(9, N, dtype=torch.float64, device="cuda")
    g = torch.zeros(P, dtype=torch.float64, device="cuda")
    pot.compute(sus2sh.box_struct(cell), t, x, e, f, v, param_grad=g)
    torch.cuda.synchronize()
    got = g.cpu().numpy()
    err = np.abs(got - pg_ref).max()
    scale = np.abs(pg_ref).max()
    print(f"  max|d(dE/dtheta)|={err:.3e} (max |dE/dtheta|={scale:.3e})", flush=True)
    assert err <= (1e-4 if fp32 else 1e-9) * max(1.0, scale)
    assert np.abs(e.cpu().numpy() - ref[0]).max() < (1e-4 if fp32 else 1e-9)
    pot.close()


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("fp32", [False, True])
def test_lane_program_matches_oracle(make_model, name, fp32):
    """Graph variant 3 (one atom per lane, wave-uniform program) on every case; a graph
    whose per-lane state does not fit the LDS of a 64-atom workgroup (l4k4 in FP64, and the
    full-l body-6 graph l2k3, whose 259 layer-1 operands of layer-2 tensors need persistent
    slots) has no lane program and is rejected cleanly."""
    case = CASES[name]
    path = make_model(name, **case["model"])
    S = case["model"].get("species_count", 2)
    pos, types, cell = _structure(case["struct"], S)
    opts = [f"T{s}" for s in range(S)] + ["sus2_float=%d" % (1 if fp32 else 0),
                                          "sus2_sh_graph_variant=3"]
    # largest per-lane slot budget whose 64-atom workgroup fits 160 KiB of LDS
    feasible = sus2sh.lane_check(path, max_slots=640 if fp32 else 320, waves=4)[0]["ok"]
    if not feasible:
        with pytest.raises(sus2sh.Sus2shError) as err:
            _run(path, cell, types, pos, options=opts)
        assert "no lane program" in err.value.message
        return
    ref = Oracle(path).compute(cell, types, pos)
    got = _run(path, cell, types, pos, options=opts)
    _compare(ref, got, fp32)


@pytest.mark.parametrize("waves", [1, 2, 4, 8])
def test_lane_program_wave_splits(make_model, waves):
    """The lane kernel with 1, 2, 4 and 8 waves per 64-atom tile (rows of every segment split
    between the waves) gives the oracle's result in both precisions; N = 500 is not a
    multiple of 64."""
    path = make_model("l3k3_fcc", **CASES["l3k3_fcc"]["model"])
    pos, types, cell = _structure(CASES["l3k3_fcc"]["struct"], 2)
    ref = Oracle(path).compute(cell, types, pos)
    for fp32 in (0, 1):
        opts = ["Cu", "Zr", f"sus2_float={fp32}", "sus2_sh_graph_variant=3",
                f"sus2_sh_lane_waves={waves}"]
        _compare(ref, _run(path, cell, types, pos, options=opts), bool(fp32))


@pytest.mark.parametrize("fp32", [False, True])
def test_force_tiles_with_neighbour_windows(make_model, fp32):
    """A periodic box of 6x6x6 cells and 2,048 atoms: eight 256-atom force tiles whose
    neighbours are summed in LDS over index windows (wrapped rows included) -- against the
    oracle, and against the same evaluation as one structure of a batch (a batch tile's window
    is the sorted range of the structures it holds)."""
    path = make_model("l3k3_fcc", **CASES["l3k3_fcc"]["model"])
    pos, types, cell = st.fcc_alloy(8, 8, 8, 3.85, [0.5, 0.5], 0.08, 31, 32)
    ref = Oracle(path).compute(cell, types, pos)
    opts = ["Cu", "Zr", f"sus2_float={int(fp32)}"]
    got = _run(path, cell, types, pos, options=opts)
    _compare(ref, got, fp32)
    import torch
    n = len(types)
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


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("tile", [1, 4, 16])
def test_graph_kernel_variants(make_model, variant, tile):
    """Every graph-kernel variant and tile the autotuner may pick gives the oracle's result."""
    path = make_model("l3k3_fcc", **CASES["l3k3_fcc"]["model"])
    pos, types, cell = _structure(CASES["l3k3_fcc"]["struct"], 2)
    ref = Oracle(path).compute(cell, types, pos)
    for fp32 in (0, 1):
        opts = ["Cu", "Zr", f"sus2_float={fp32}", f"sus2_sh_graph_variant={variant}",
                f"sus2_sh_graph_tile={tile}"]
        if tile == 16 and not fp32:
            # 2 x 756 slots x 16 atoms x 8 B exceeds the LDS of a workgroup: rejected cleanly
            with pytest.raises(sus2sh.Sus2shError) as err:
                _run(path, cell, types, pos, options=opts)
            assert "too large" in err.value.message
            continue
        got = _run(path, cell, types, pos, options=opts)
        _compare(ref, got, bool(fp32))


@pytest.mark.parametrize("fp32", [False, True])
def test_batch_windows_mixed_sizes(make_model, fp32):
    """A batch of small cells (a 128-atom force tile spans several of them: one window, the
    sorted range of its structures) and a 2,916-atom cell (its range exceeds the 2,048-atom
    window capacity: its tiles fall back to per-edge atomics), each against the oracle."""
    import torch
    path = make_model("l3k3_fcc", **CASES["l3k3_fcc"]["model"])
    structs = [st.fcc_alloy(2, 2, 2, 3.85, [0.5, 0.5], 0.08, 61, 62),
               st.fcc_alloy(3, 3, 3, 3.85, [0.5, 0.5], 0.08, 63, 64),
               st.fcc_alloy(9, 9, 9, 3.85, [0.5, 0.5], 0.08, 65, 66),
               st.fcc_alloy(4, 4, 4, 3.85, [0.5, 0.5], 0.08, 67, 68),
               st.fcc_alloy(3, 3, 2, 3.85, [0.5, 0.5], 0.08, 69, 70)]
    sizes = [len(t) for _, t, _ in structs]
    assert sizes[2] == 2916 and max(s for k, s in enumerate(sizes) if k != 2) <= 256
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    n = int(offs[-1])
    print(f"[case] batch windows, sizes {sizes}, fp32={fp32}", flush=True)
    pot = sus2sh.SUS2_SH(path, n, ["Cu", "Zr", f"sus2_float={int(fp32)}"])
    dev = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a, dtype=dt), device="cuda")
    e = torch.zeros(n, dtype=torch.float64, device="cuda")
    f = torch.zeros(3, n, dtype=torch.float64, device="cuda")
    v = torch.zeros(9, n, dtype=torch.float64, device="cuda")
    sus2sh.compute_batch(pot, [c for _, _, c in structs], offs,
                         dev(np.concatenate([t for _, t, _ in structs]), np.int32),
                         dev(np.concatenate([p for p, _, _ in structs], axis=1), np.float64),
                         e, f, v)
    torch.cuda.synchronize()
    pot.close()
    e, f, v = e.cpu().numpy(), f.cpu().numpy(), v.cpu().numpy()
    orc = Oracle(path)
    for k, (pos, types, cell) in enumerate(structs):
        sl = slice(offs[k], offs[k + 1])
        _compare(orc.compute(cell, types, pos), (e[sl], f[:, sl], v[:, sl]), fp32)
