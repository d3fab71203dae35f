import importlib.util
import pathlib
import sys

ROOT = pathlib.Path(__file__).resolve().parents[1]
PKG_DIR = ROOT / "seiswavelab--curvilinear_amd"
if "seiswavelab_curvilinear_amd" not in sys.modules:
    _spec = importlib.util.spec_from_file_location("seiswavelab_curvilinear_amd", PKG_DIR / "__init__.py",
                                                   submodule_search_locations=[str(PKG_DIR)])
    _mod = importlib.util.module_from_spec(_spec)
    sys.modules[_spec.name] = _mod
    _spec.loader.exec_module(_mod)
sys.path.insert(0, str(pathlib.Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
