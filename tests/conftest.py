import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gpumd-sus2-sh_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: test needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def model_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("models")


def _model(model_dir, name, **kw):
    import sus2sh
    path = os.path.join(str(model_dir), name + ".mtp")
    if not os.path.exists(path):
        spec = sus2sh.synthetic_model(**kw)
        sus2sh.write_mtp(spec, path)
    return path


@pytest.fixture(scope="session")
def make_model(model_dir):
    def make(name, **kw):
        return _model(model_dir, name, **kw)
    return make
