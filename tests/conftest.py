import os
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
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:  # noqa: BLE001
        return False


@pytest.fixture(scope='session')
def need_gpu():
    if not gpu_available():
        pytest.fail('GPU test collected without a visible GPU')
