import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT/'tests'/'golden'


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a ROCm GPU (gfx950) and the built HIP library')


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:   # noqa: BLE001
        return False


@pytest.fixture
def gpu():
    if not gpu_available():
        pytest.fail('this test needs a ROCm GPU')
    import torch
    return torch.device('cuda:0')
