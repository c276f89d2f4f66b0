import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# CATN_EMU_LIB selects another emulator build for the "emu" cases, e.g. the
# multi-threaded lane emulator (tests/emul/catn_emul_mt.cpp) under a sanitizer.
EMUL_LIB = os.environ.get("CATN_EMU_LIB") or os.path.join(ROOT, "tests", "emul", "libcatn_emul.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def catn():
    import __graft_entry__ as G
    return G.load_package()


@pytest.fixture(scope="session")
def emu_ctx(catn):
    """CPU emulator build of the engine (test-only library)."""
    if not os.path.exists(EMUL_LIB):
        pytest.skip("emulator not built (run __graft_entry__.build())")
    return catn.Context(lib=EMUL_LIB)


@pytest.fixture(scope="session")
def gpu_ctx(catn):
    """The product library on cuda:0; fails loudly when it cannot load."""
    return catn.Context(0)
