import sys
from pathlib import Path

import pytest

TESTS = Path(__file__).resolve().parent
if str(TESTS) not in sys.path:
    sys.path.insert(0, str(TESTS))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")
    # torch ships its own HIP runtime under the same soname as the system one;
    # whichever loads first serves the process, and torch only finds the GPU
    # when it is its own.  So when GPU tests may run, torch is imported before
    # libscr.so is loaded (the torch harness tests need it; the others run on
    # that runtime too).
    if "not gpu" not in (config.option.markexpr or ""):
        try:
            import torch  # noqa: F401
        except ImportError:
            pass


@pytest.fixture(scope="session")
def ctx():
    """One GPU context for the whole session (fails loudly without a GPU)."""
    from _pkg import scr
    c = scr.Context(0)
    yield c
    c.close()
