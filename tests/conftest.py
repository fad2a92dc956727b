import io
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def has_gpu():
    try:
        import thousandthchicken_amd as tca
        n = __import__("ctypes").c_int(0)
        return tca.lib().tc_device_count(__import__("ctypes").byref(n)) == 0 and n.value > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu_decoder():
    import thousandthchicken_amd as tca
    dec = tca.Decoder()
    yield dec
    dec.close()
