import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def corecheck():
    """The solver's shared core compiled for the host (tests/native/core_check.cpp)."""
    import ctypes
    out = os.path.join(ROOT, "tests", "native", "_build", "libcorecheck.so")
    src = os.path.join(ROOT, "tests", "native", "core_check.cpp")
    hdr = [os.path.join(ROOT, "doublyperiodicstokes_amd", "csrc", h)
           for h in ("dps_core.h", "dps_fft.h", "dps_seg.h", "dps_tables.h")]
    if not os.path.exists(out) or any(os.path.getmtime(f) > os.path.getmtime(out) for f in [src] + hdr):
        os.makedirs(os.path.dirname(out), exist_ok=True)
        subprocess.check_call(["g++", "-O2", "-std=c++20", "-shared", "-fPIC", "-pthread",
                               "-I", os.path.join(ROOT, "doublyperiodicstokes_amd", "csrc"),
                               "-o", out, src])
    return ctypes.CDLL(out)


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
