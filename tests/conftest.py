import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    # Build what is missing (the built libraries are used as they are otherwise).
    need = [os.path.join(ROOT, "oracle", "liboracle.so"),
            os.path.join(ROOT, "flashraw_amd", "lib", "libflashraw_hip.so"),
            os.path.join(ROOT, "flashraw_amd", "lib", "libfrsynth.so")]
    if not all(os.path.exists(p) for p in need):
        import importlib
        entry = importlib.import_module("__graft_entry__")
        entry.build()


@pytest.fixture(scope="session")
def ctx():
    import flashraw_amd as F
    if not F.available():
        pytest.skip("no HIP device")
    return F.context()
