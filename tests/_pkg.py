"""Import helpers for the tests: the product package (directory name has a
hyphen, so it is loaded by path) and the CPU oracle."""
import importlib.util
import sys
from pathlib import Path

REPO = Path(__file__).resolve().parents[1]
PKG_DIR = REPO / "satin-computerasteriser_amd"


def _load_pkg():
    name = "satin_computerasteriser_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def _load_oracle():
    if str(REPO) not in sys.path:
        sys.path.insert(0, str(REPO))
    from oracle import oracle as orc
    return orc


scr = _load_pkg()
orc = _load_oracle()
