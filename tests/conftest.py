import ctypes
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / 'bm25-npu_amd'):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs through libbm25mi on cuda:0)')
    config.addinivalue_line('markers', 'slow: larger CPU cases')


def gpu_available():
    """Whether HIP sees a GPU, asked of the ROCm HIP runtime the library links (libamdhip64.so.7)
    rather than of torch: torch's wheel bundles a HIP runtime of its own, and a process that
    initialises that one first and loads the library's afterwards gets 'no ROCm-capable device'
    from every call of the second."""
    try:
        hip = ctypes.CDLL('libamdhip64.so.7')
    except OSError:
        return False
    count = ctypes.c_int(0)
    return hip.hipGetDeviceCount(ctypes.byref(count)) == 0 and count.value > 0


@pytest.fixture(scope='session')
def need_gpu():
    if not gpu_available():
        pytest.fail('GPU test collected without a visible GPU')
