"""Host-side `.mtp` loader, validation errors and the C ABI surface (CPU only).

The validation list follows load_sh_model / build_explicit_sh_graph_metadata
(src/force/sus2_sh.cu:2105-2482, 1913-2093); each negative case must come back as
SUS2SH_ERR_INPUT with the reference's message, never as a process exit.
"""
import ctypes
import os
import re

import pytest

import sus2sh
from sus2sh import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "sus2sh.h")).read()
    declared = set(re.findall(r"\b(sus2sh_[a-z_]+)\s*\(", header))
    assert declared == set(_capi.EXPORTED_SYMBOLS)
    lib = ctypes.CDLL(_capi.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name


@pytest.fixture(scope="module")
def good(make_model):
    return make_model("loader_l2k2", l_max=2, k_max=2, body_order=5, rb_size=8, max_dist=5.0,
                      seed=7)


def test_round_trip_and_probe(good, tmp_path):
    d = sus2sh.model_check(good)
    assert d["standard_graph"] == 1
    assert (d["moments_count"], d["basic_count"], d["product_count"], d["scalar_count"]) == \
        (210, 18, 535, 91)
    assert d["row_term_count"] == d["product_count"]
    assert d["back_term_count"] == 2 * d["product_count"]
    assert sus2sh.probe(good)
    other = tmp_path / "other.mtp"
    other.write_text(open(good).read().replace("potential_tag = SUS2-SH", "potential_tag = SUS2"))
    assert not sus2sh.probe(other)
    notmtp = tmp_path / "nep.txt"
    notmtp.write_text("nep4 2 Cu Zr\n" + open(good).read())
    assert not sus2sh.probe(notmtp)
    assert not sus2sh.probe(tmp_path / "missing.mtp")


def test_non_standard_graph_is_accepted(good, tmp_path):
    """A valid explicit graph that differs from the standard one loads (flag off)."""
    text = open(good).read()
    m = re.search(r"sh_products = \{\{(\d+), (\d+), (\d+), ([^}]+)\}", text)
    coeff = float(m.group(4))
    text2 = text.replace(m.group(0), "sh_products = {{%s, %s, %s, %r}" % (
        m.group(1), m.group(2), m.group(3), coeff * 1.5), 1)
    p = tmp_path / "nonstd.mtp"
    p.write_text(text2)
    assert sus2sh.model_check(p)["standard_graph"] == 0


def _mutate(text, old, new, count=1):
    assert old in text, old
    return text.replace(old, new, count)


CASES = [
    ("version", lambda t: _mutate(t, "version = 1.1.0", "version = 1.0.0"),
     "supports only version = 1.1.0"),
    ("tag", lambda t: _mutate(t, "potential_tag = SUS2-SH", "potential_tag = SUS2-X"),
     "non-SUS2-SH model"),
    ("scaling_map", lambda t: _mutate(t, "scaling_map = LK", "scaling_map = K"),
     "requires scaling_map = LK"),
    ("lmax", lambda t: _mutate(t, "sh_l_max = 2", "sh_l_max = 7"), "sh_l_max in [0,6]"),
    ("body", lambda t: _mutate(t, "sh_body_order = 5", "sh_body_order = 7"),
     "sh_body_order in [2,6]"),
    ("mr", lambda t: _mutate(t, "radial_funcs_count = 6", "radial_funcs_count = 5"),
     "inconsistent sh_k_max/radial_funcs_count"),
    ("rb", lambda t: _mutate(t, "radial_basis_size = 8", "radial_basis_size = 17"),
     "rb_size<=16"),
    ("parity", lambda t: _mutate(t, "sh_parity = even", "sh_parity = odd"), "even parity"),
    ("radial_type", lambda t: _mutate(t, "RBChebyshev_sss", "RBBessel"),
     "Unsupported SUS2_SH radial_basis_type"),
    ("missing", lambda t: _mutate(t, "alpha_scalar_moments =", "alpha_scalar_momentz ="),
     "Missing token in SUS2-SH model file: alpha_scalar_moments ="),
    ("body_lmax_size", lambda t: _mutate(t, "sh_body_l_max = {2, 2, 2, 2}",
                                         "sh_body_l_max = {2, 2, 2}"),
     "Unexpected SUS2_SH sh_body_l_max size"),
    ("body_lmax_range", lambda t: _mutate(t, "sh_body_l_max = {2, 2, 2, 2}",
                                          "sh_body_l_max = {2, 2, 3, 2}"),
     "sh_body_l_max entry is out of range"),
    ("basic_entry", lambda t: _mutate(t, "alpha_index_basic = {{1, 2, -2}",
                                      "alpha_index_basic = {{1, 2, -3}"),
     "Invalid SUS2_SH alpha_index_basic entry"),
    ("product_count", lambda t: _mutate(t, "sh_product_count = 535", "sh_product_count = 534"),
     "Unexpected SUS2_SH sh_products size"),
    ("product_range", lambda t: re.sub(r"sh_products = \{\{\d+,", "sh_products = {{999,", t, 1),
     "SUS2_SH product index out of range"),
    ("product_into_basic", lambda t: re.sub(r"sh_products = \{\{(\d+), (\d+), \d+,",
                                            r"sh_products = {{\1, \2, 3,", t, 1),
     "writes into a basic moment"),
    ("coeff_dims", lambda t: re.sub(r"shift_coeffs = \{([^,]+),", r"shift_coeffs = {\1, 0.0,", t, 1),
     "Unexpected SUS2_SH coefficient dimensions"),
    ("radial_row", lambda t: re.sub(r"\t\t\{([^,]+),", r"\t\t{\1, 0.0,", t, 1),
     "Unexpected SUS2_SH radial_coeffs row size"),
    ("mapping", lambda t: re.sub(r"alpha_moment_mapping = \{(\d+),", r"alpha_moment_mapping = {209,", t, 1),
     None),
]


@pytest.mark.parametrize("name,mutate,msg", CASES, ids=[c[0] for c in CASES])
def test_loader_rejects_invalid_models(good, tmp_path, name, mutate, msg):
    p = tmp_path / f"{name}.mtp"
    p.write_text(mutate(open(good).read()))
    if msg is None:
        # node 209 is defined (last product target), so this is a valid but non-standard model
        assert sus2sh.model_check(p)["standard_graph"] == 0
        return
    with pytest.raises(sus2sh.Sus2shError) as err:
        sus2sh.model_check(p)
    assert err.value.code == _capi.SUS2SH_ERR_INPUT
    assert msg in err.value.message


def test_topological_order_and_undefined_moment(good, tmp_path):
    text = open(good).read()
    prods = re.search(r"sh_products = \{(.*)\}\n", text).group(1)
    groups = re.findall(r"\{[^{}]*\}", prods)
    # move the last product first: it consumes nodes defined later -> not topological
    reordered = [groups[-1]] + groups[:-1]
    bad = text.replace(prods, ", ".join(reordered))
    p = tmp_path / "topo.mtp"
    p.write_text(bad)
    with pytest.raises(sus2sh.Sus2shError) as err:
        sus2sh.model_check(p)
    assert "not topologically ordered" in err.value.message or "undefined moment" in err.value.message


def test_create_reports_errors_without_exiting(tmp_path):
    bad = tmp_path / "none.mtp"
    with pytest.raises(sus2sh.Sus2shError) as err:
        sus2sh.model_check(bad)
    assert "Failed to open SUS2-SH model file" in err.value.message


@pytest.mark.parametrize("opt", ["sus2_sh_basic_shape=96x2", "sus2_sh_force_shape=256x3",
                                 "sus2_sh_force_shape=256", "sus2_sh_basic_shape=512x1",
                                 "sus2_sh_basic_mfma=fp16"])
def test_invalid_edge_shape_option_is_an_input_error(good, opt):
    """Edge-kernel shapes outside threads {64, 128, 256} x lanes per atom {1, 2, 4, 8} are
    rejected while the options are parsed (before any device call) with a message naming the
    option."""
    with pytest.raises(sus2sh.Sus2shError) as err:
        sus2sh.SUS2_SH(good, 16, ["Cu", "Zr", opt])
    assert err.value.code == _capi.SUS2SH_ERR_INPUT
    assert opt.split("=")[0] in err.value.message
