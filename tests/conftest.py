import importlib.util
import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
if str(REPO / "tests") not in sys.path:
    sys.path.insert(0, str(REPO / "tests"))


def _load_pkg():
    name = "cloud_renderer_amd"
    if name in sys.modules:
        return sys.modules[name]
    pkg_dir = REPO / "cloud-renderer_amd"
    spec = importlib.util.spec_from_file_location(name, pkg_dir / "__init__.py",
                                                  submodule_search_locations=[str(pkg_dir)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


_load_pkg()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: full-size checks (minutes of CPU oracle time)")


@pytest.fixture(scope="session")
def cv():
    return _load_pkg()


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    oracle_lib.lib()
    return oracle_lib


@pytest.fixture(scope="session")
def gpu_ctx(cv):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    ctx = cv.Context(0)
    yield ctx
