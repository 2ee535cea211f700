"""Block program -- the tensor-organised program of the block graph kernel (graph variant 4) --
executed on the host through the C ABI (sus2sh_block_check) against the FP64 oracle's flat
product graph (oracle_graph: src/force/sus2_sh.cu:4799-4846) on seeded basic vectors.

Bar: E' and dE'/dbasic agree to 1e-11 relative to their scale with FP64 coefficients (the
program sums in another order: per tensor K' v, E' = 1/2 sum v . (g + d), reverse through the
CG blocks), and to 1e-6 with the coefficients rounded to FP32 (the device words of the FP32
kernel).  The compile-time coupling table the kernel unrolls (csrc/cg_table.hpp, written by
csrc/gen_cg_table.py) is pinned against the library's real_cg, and the generator reproduces it.
"""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import sus2sh
from oracle.binding import Oracle

CSRC = os.path.join(os.path.dirname(__file__), "..", "gpumd-sus2-sh_amd", "csrc")

STANDARD = [
    # (l_max, k_max, body_order, body_l_max)
    (0, 1, 2, None),
    (0, 3, 3, None),
    (1, 2, 4, None),
    (2, 2, 5, None),
    (2, 3, 4, None),
    (3, 3, 5, (3, 3, 3, 3)),
    (3, 2, 5, (3, 2, 1, 0)),
    (3, 1, 3, None),
]


def _check(path, n=4, seed=0):
    orc = Oracle(path)
    B = orc.dims["B"]
    rng = np.random.default_rng(seed)
    basics = rng.normal(0.0, 0.4, size=(n, B))
    e0, g0 = orc.graph(basics)
    scale = max(1.0, np.abs(e0).max(), np.abs(g0).max())
    info = None
    for waves in (4, 8, 16):
        for rb, tol in ((8, 1e-11), (4, 1e-6)):
            info, e, g = sus2sh.block_check(path, basics, waves=waves, real_bytes=rb)
            print(f"[block] {os.path.basename(path)} waves={waves} real_bytes={rb}: {info}",
                  flush=True)
            assert info["ok"] == 1
            assert np.abs(e - e0).max() <= tol * scale
            assert np.abs(g - g0).max() <= tol * scale
    # a slot target one below the whole-class program: the classes of degree >= 2 above it
    # are halved by components when that reaches the target (else the program is unchanged)
    hinfo, e, g = sus2sh.block_check(path, basics, target_slots=info["slots"] - 1, real_bytes=8)
    print(f"[block] halved: {hinfo}", flush=True)
    assert hinfo["ok"] == 1 and hinfo["units"] >= info["units"]
    assert (hinfo["slots"] < info["slots"]) == (hinfo["units"] > info["units"])
    assert np.abs(e - e0).max() <= 1e-11 * scale and np.abs(g - g0).max() <= 1e-11 * scale
    return info


@pytest.mark.parametrize("L,K,body,blm", STANDARD)
def test_standard_graphs(make_model, L, K, body, blm):
    path = make_model(f"block_l{L}k{K}b{body}_{blm}", l_max=L, k_max=K, body_order=body,
                      body_l_max=blm, rb_size=4, max_dist=4.0, seed=200 + 7 * L + K + body,
                      species_count=2)
    _check(path)


def test_c3_graph_counts(make_model):
    """The l3k3 body-5 graph of the benchmark: 136 layer-1 CG blocks (3,332 forward terms,
    twice that in reverse), 145 quadratic-form tensors in 4 degree batches, 258 Real slots;
    halved to 210 slots for two FP32 workgroups per CU."""
    path = make_model("block_l3k3_c3", l_max=3, k_max=3, body_order=5, rb_size=8, max_dist=5.0,
                      seed=3003, scaling=0.5)
    info = _check(path, n=3)
    assert (info["fwd_blocks"], info["fwd_terms"], info["bwd_terms"]) == (136, 3332, 6664)
    assert (info["tensors"], info["batches"], info["slots"]) == (145, 4, 258)
    # two FP32 workgroups per CU: 48 x 64 doubles of adjoints + 224 x 64 floats = 80 KiB; the
    # degree-2 class (42 tensors, 210 values) is halved into 126 + 84 values
    h = sus2sh.block_check(path, target_slots=224, real_bytes=4)[0]
    assert (h["ok"], h["slots"], h["units"], h["batches"], h["fwd_blocks"]) == (1, 210, 8, 5, 178)


def test_pruned_explicit_graph(tmp_path):
    """A pruned explicit graph (body orders <= 3 and every other higher-order scalar kept,
    sus2sh.subgraph) keeps only the live tensors and matches the oracle."""
    import copy
    spec = sus2sh.synthetic_model(3, 2, 5, rb_size=4, max_dist=4.0, seed=31)
    g = spec.graph
    keep = (g.scalar_body <= 3) | (np.arange(g.scalar_count) % 2 == 0)
    pruned = copy.deepcopy(spec)
    pruned.graph = sus2sh.subgraph(g, keep)
    pruned.moment_coeffs = spec.moment_coeffs[keep]
    path = sus2sh.write_mtp(pruned, str(tmp_path / "pruned.mtp"))
    assert sus2sh.model_check(path)["standard_graph"] == 0
    _check(path)


@pytest.mark.parametrize("L,K,body,why", [(4, 2, 4, "degree above 3"),
                                          (2, 3, 6, "deeper than two layers")])
def test_unsupported_models(make_model, L, K, body, why):
    path = make_model(f"block_no_l{L}k{K}b{body}", l_max=L, k_max=K, body_order=body,
                      rb_size=4, max_dist=4.0, seed=40 + L)
    info, e, g = sus2sh.block_check(path, np.zeros((1, sus2sh.model_check(path)["basic_count"])))
    assert info["ok"] == 0 and e is None and g is None


def test_cg_table_matches_real_cg():
    txt = open(os.path.join(CSRC, "cg_table.hpp")).read()
    arr = lambda name: [int(v) for v in re.search(name + r"\[kCgTypes\] = \{(.*)\}", txt).group(1).split(",")]
    L1, L2, LL = arr("kCgL1"), arr("kCgL2"), arr("kCgLL")
    blocks = re.findall(r"struct CgTable<(\d+)> \{.*?e\[n\] = \{(.*?)\};", txt, re.S)
    assert len(blocks) == len(L1) == 44
    for t, body in blocks:
        t = int(t)
        ent = {(int(a), int(b), int(c)): float(v)
               for a, b, c, v in re.findall(r"\{(\d+), (\d+), (\d+), ([-0-9.e]+)\}", body)}
        l1, l2, L = L1[t], L2[t], LL[t]
        for M in range(2 * L + 1):
            for m1 in range(2 * l1 + 1):
                for m2 in range(2 * l2 + 1):
                    r = sus2sh.real_cg(l1, m1 - l1, l2, m2 - l2, L, M - L)
                    assert ent.get((M, m1, m2), 0.0) == pytest.approx(r, abs=1e-15)
                    assert abs(r) <= 1e-13 or (M, m1, m2) in ent


def test_cg_table_generator_reproduces_header(tmp_path):
    out = tmp_path / "cg_table.hpp"
    subprocess.run([sys.executable, os.path.join(CSRC, "gen_cg_table.py"), str(out)], check=True)
    assert out.read_text() == open(os.path.join(CSRC, "cg_table.hpp")).read()
