"""Test configuration: the `gpu` marker and the loader of the package
directory `visual-cortex_amd/` (not an importable name) as
`visual_cortex_amd`."""
import importlib.util
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG_DIR = ROOT / "visual-cortex_amd"


def _load_pkg():
    name = "visual_cortex_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, PKG_DIR / "__init__.py", submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


_load_pkg()
if str(ROOT / "oracle") not in sys.path:
    sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def pkg():
    return _load_pkg()


@pytest.fixture(scope="session")
def orc(pkg):
    import oracle
    from visual_cortex_amd import abi
    return oracle.Oracle(abi)
