"""GPUMD input files (run.in potential line, extended XYZ) -- CPU only, on the reference's
own smoke inputs (tests/golden/gpumd_smoke) and synthetic variants."""
import os

import numpy as np
import pytest

import sus2sh
from sus2sh import gpumd_io

GOLD = os.path.join(os.path.dirname(__file__), "golden", "gpumd_smoke")
SYMBOLS = ["H", "C", "N", "O", "Cl", "K", "Co", "Ni"]


@pytest.fixture(scope="module")
def eight_species(model_dir):
    spec = sus2sh.synthetic_model(2, 2, 5, species_count=8, rb_size=6, max_dist=5.0, seed=81,
                                  radial_basis_type="RBLaguerre_log1p", scaling=0.5)
    return sus2sh.write_mtp(spec, os.path.join(str(model_dir), "smoke_lag8.mtp"))


def test_reference_smoke_xyz():
    xyz = gpumd_io.read_xyz(os.path.join(GOLD, "model.xyz"))
    assert xyz["symbols"] == ["H", "C"]
    assert xyz["pbc"] == (1, 1, 1)
    assert np.array_equal(xyz["cell"], np.diag([20.0, 20.0, 20.0]))
    assert np.array_equal(xyz["positions"], np.array([[5.0, 7.0], [5.0, 5.0], [5.0, 5.0]]))


def test_reference_smoke_potential_line(eight_species):
    path, symbols, options = gpumd_io.potential_line(os.path.join(GOLD, "run.in"), eight_species)
    assert path == eight_species
    assert symbols == SYMBOLS
    assert options == ["sus2_float=1", "sus2_radial_direct=1"]
    _, opts, cell, pbc, types, pos = gpumd_io.gpumd_inputs(
        os.path.join(GOLD, "run.in"), os.path.join(GOLD, "model.xyz"), eight_species)
    assert opts == SYMBOLS + options
    assert list(types) == [0, 1] and types.dtype == np.int32
    assert sus2sh.model_check(path)["species_count"] == 8


def test_lattice_columns_pbc_and_extra_properties(tmp_path):
    p = tmp_path / "tri.xyz"
    p.write_text('3\nLattice="10 0 0 1 9 0 0.5 0.2 8" pbc="T F T" '
                 'Properties=species:S:1:pos:R:3:mass:R:1:vel:R:3 energy=-3.0\n'
                 "Cu 0 0 0 63.5 0.1 0 0\nZr 1 2 3 91.2 0 0.2 0\nCu 4 5 6 63.5 0 0 0.3\n")
    xyz = gpumd_io.read_xyz(str(p))
    assert np.array_equal(xyz["cell"], np.array([[10, 0, 0], [1, 9, 0], [0.5, 0.2, 8.0]]))
    assert xyz["pbc"] == (1, 0, 1)
    assert np.allclose(xyz["masses"], [63.5, 91.2, 63.5])
    assert np.allclose(xyz["velocities"], np.diag([0.1, 0.2, 0.3]))
    assert np.allclose(xyz["positions"][:, 1], [1, 2, 3])


@pytest.mark.parametrize("text,message", [
    ("2 3\nLattice=\"1 0 0 0 1 0 0 0 1\" Properties=species:S:1:pos:R:3\nH 0 0 0\nH 1 1 1\n",
     "one value"),
    ("1\nLattice=\"1 0 0 0 1 0 0 0 1\" Properties=species:S:1:pos:R:3\nH 0 0 0\n", ">= 2"),
    ("2\nProperties=species:S:1:pos:R:3\nH 0 0 0\nH 1 1 1\n", "lattice"),
    ("2\nLattice=\"1 0 0 0 1 0 0 0 1\"\nH 0 0 0\nH 1 1 1\n", "properties"),
    ("2\nLattice=\"1 0 0 0 1 0 0 0 1\" pbc=\"T X T\" Properties=species:S:1:pos:R:3\n"
     "H 0 0 0\nH 1 1 1\n", "T/F"),
])
def test_xyz_errors(tmp_path, text, message):
    p = tmp_path / "bad.xyz"
    p.write_text(text)
    with pytest.raises(ValueError) as err:
        gpumd_io.read_xyz(str(p))
    assert message in str(err.value)


def test_potential_line_errors(tmp_path, eight_species):
    r = tmp_path / "run.in"
    r.write_text("velocity 300\n")
    with pytest.raises(ValueError, match="no 'potential'"):
        gpumd_io.potential_line(str(r), eight_species)
    r.write_text(f"potential {eight_species} H C N\n")
    with pytest.raises(ValueError, match="fewer element symbols"):
        gpumd_io.potential_line(str(r))
    x = tmp_path / "m.xyz"
    x.write_text('2\nLattice="9 0 0 0 9 0 0 0 9" Properties=species:S:1:pos:R:3\nH 0 0 0\nAu 1 1 1\n')
    r.write_text(f"potential {eight_species} " + " ".join(SYMBOLS) + "\n")
    with pytest.raises(ValueError, match="not on the potential line"):
        gpumd_io.gpumd_inputs(str(r), str(x))
