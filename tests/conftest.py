import sys
from pathlib import Path

import pytest

# PyTorch (device buffers in the device-array tests) ships its own HIP runtime and exports it process-wide. When both
# are in one process it must be loaded before libdecaes_amd.so, whose HIP calls then bind to that same runtime; a
# second runtime initialized after torch's finds no device.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / "decaes.jl_amd", ROOT / "tests", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; parity tests through the C ABI")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib

    oracle_lib.build()
    return oracle_lib


@pytest.fixture(scope="session")
def dmi():
    import decaes_amd

    decaes_amd.load_library()
    return decaes_amd
