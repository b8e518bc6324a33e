import importlib.util
import pathlib
import subprocess
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a visible MI355X (runs the HIP path)")


def load_package():
    if "prop_streams2_amd" in sys.modules:
        return sys.modules["prop_streams2_amd"]
    pkg_dir = ROOT / "prop-streams2_amd"
    spec = importlib.util.spec_from_file_location("prop_streams2_amd", pkg_dir / "__init__.py",
                                                  submodule_search_locations=[str(pkg_dir)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["prop_streams2_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


PKG = load_package()
if str(ROOT / "tests") not in sys.path:
    sys.path.insert(0, str(ROOT / "tests"))


@pytest.fixture(scope="session")
def pkg():
    return PKG


@pytest.fixture(scope="session")
def oracle():
    import oracle_backend
    return oracle_backend.load_oracle()


@pytest.fixture(scope="session")
def h2_tables():
    return PKG.cases.h2_air_mechanism().tables()
