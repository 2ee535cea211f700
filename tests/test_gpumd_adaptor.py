"""The GPUMD C++ adaptor (include/sus2sh.hpp) compiles against GPUMD-shaped Box / GPU_Vector
stand-ins and works through the C ABI without a GPU: the probe recognises a model file
(src/force/sus2_sh.cu:7771-7780) and the constructor turns the loader's and the device's
errors into sus2sh::Error with the reference's message text instead of exit(1)
(src/force/sus2_sh.cu:269-273).  The program is tests/cpp/gpumd_adaptor_mock.cpp."""
import os
import shutil
import subprocess

import pytest

import sus2sh

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def adaptor_exe(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    if not os.path.exists(sus2sh.LIB_PATH):
        pytest.fail(f"{sus2sh.LIB_PATH} is not built")
    out = tmp_path_factory.mktemp("adaptor") / "gpumd_adaptor_mock"
    libdir = os.path.dirname(sus2sh.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
           "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "cpp", "gpumd_adaptor_mock.cpp"),
           "-o", str(out), "-L", libdir, "-lsus2sh", f"-Wl,-rpath,{libdir}",
           "-Wl,-rpath-link,/opt/rocm/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return out


def test_adaptor_compiles_and_reports_errors(adaptor_exe, tmp_path):
    spec = sus2sh.synthetic_model(2, 2, 3, species_count=2, rb_size=6, max_dist=4.5, seed=11)
    good = sus2sh.write_mtp(spec, str(tmp_path / "good.mtp"))
    text = open(good).read()
    bad = tmp_path / "bad.mtp"
    bad.write_text(text.replace("sh_l_max = 2", "sh_l_max = 9"))
    other = tmp_path / "other.txt"
    other.write_text("not a potential\n")
    res = subprocess.run([str(adaptor_exe), good, str(bad), str(other)], capture_output=True,
                         text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "probe model=1 not_a_model=0" in res.stdout
    assert "invalid model: code=" in res.stdout
    assert "gpumd compute form instantiated=1" in res.stdout
