"""The drop-in stage classes (CoefficientCoder -> Quantizer -> DWT ->
Preprocessor, run in the reference's order) give the same image as the fused
path and the oracle."""
import os

import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# staged decode rejects sub-sampled components
NAMES = sorted(f for f in os.listdir(GOLDEN) if f.endswith(".j2k") and not f.startswith("pkg_sub_"))


@pytest.mark.parametrize("name", NAMES)
def test_staged_equals_fused_and_oracle(gpu_decoder, name):
    cs = open(os.path.join(GOLDEN, name), "rb").read()
    print(f"case {name}", flush=True)
    fused = gpu_decoder.decode(cs)
    staged = tca.decode_staged(cs)
    ref = oracle.decode(cs).transpose(1, 2, 0)
    lim = 0 if "_ll_" in name else 1
    assert np.abs(staged.astype(np.int32) - ref).max() <= lim
    assert np.abs(fused.astype(np.int32) - ref).max() <= lim
    if lim == 0:
        assert np.array_equal(staged, fused)
