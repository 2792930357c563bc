import json
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))

GOLDEN = Path(__file__).resolve().parent / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (run with -m gpu)")


def load_golden(name: str):
    return json.loads((GOLDEN / f"{name}.json").read_text())


@pytest.fixture(scope="session")
def gpu():
    """The native library with a usable GPU; GPU tests fail loudly when either is missing."""
    import samaya_amd as smy
    lib = smy.load()
    assert lib.smy_gpu_available() == 1, "no usable gfx950 GPU visible to the native library"
    return lib
