"""The oracle's parameter adjoint dE/dtheta against central finite differences of the
oracle energy over rewritten model files (CPU only)."""
import os

import numpy as np
import pytest

import sus2sh
from sus2sh import structures as st
from oracle.binding import Oracle


@pytest.mark.parametrize("kind", ["RBChebyshev_sss", "RBChebyshev_sss_rational"])
def test_param_gradient_central_differences(tmp_path, kind):
    spec = sus2sh.synthetic_model(2, 2, 5, species_count=2, rb_size=5, max_dist=4.6, seed=77,
                                  scal_mode="random", radial_basis_type=kind, scaling=0.5)
    base = str(tmp_path / "base.mtp")
    sus2sh.write_mtp(spec, base)
    pos, types, cell = st.fcc_alloy(2, 2, 2, 3.6, [0.5, 0.5], 0.1, 1, 2)
    e, f, v, pg = Oracle(base).compute(cell, types, pos, param_grad=True)
    theta = sus2sh.param_vector(spec)
    assert len(theta) == len(pg)
    rng = np.random.default_rng(0)
    idx = list(range(0, 4)) + list(rng.choice(np.arange(4, len(theta)), 30, replace=False))
    h = 1e-5
    for k in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        pp = str(tmp_path / "p.mtp")
        pm = str(tmp_path / "m.mtp")
        sus2sh.write_mtp(sus2sh.with_params(spec, tp), pp)
        sus2sh.write_mtp(sus2sh.with_params(spec, tm), pm)
        fd = (Oracle(pp).compute(cell, types, pos)[0].sum()
              - Oracle(pm).compute(cell, types, pos)[0].sum()) / (2 * h)
        assert fd == pytest.approx(pg[k], abs=1e-6 * max(1.0, abs(fd))), k
