"""The package's encoder (test-data generator) against the oracle and OpenJPEG:
its lossless codestreams decode back to the source, and an independent decoder
(OpenJPEG 2.5.4) reads them identically."""
import numpy as np
import pytest

from oracle import oracle
from tests import imaging
from thousandthchicken_amd import synth

CASES = [
    ((64, 64, 3, 8), dict(levels=2, cb=64)),
    ((100, 77, 3, 8), dict(levels=3, cb=32)),
    ((150, 110, 3, 12), dict(levels=4, cb=32, tile=(64, 64))),
    ((120, 90, 1, 16), dict(levels=5, cb=64)),
    ((97, 131, 1, 8), dict(levels=0, cb=32)),
    ((33, 47, 3, 10), dict(levels=6, cb=32, tile=(16, 24))),
]


@pytest.mark.parametrize("shape,kw", CASES, ids=[str(c[0]) for c in CASES])
def test_encoder_lossless_round_trip(shape, kw):
    h, w, c, bits = shape
    a = synth.image(h, w, c, 7, bits)
    cs = synth.encode(a, bits=bits, **kw)
    d = oracle.decode(cs).transpose(1, 2, 0)
    assert np.array_equal(d, a.reshape(h, w, c).astype(np.int32))
    if imaging.pillow_j2k() is not None and bits in (8, 16):
        r = imaging.decode_openjpeg(cs)
        assert np.array_equal(r.reshape(h, w, c), a.reshape(h, w, c))


@pytest.mark.parametrize("bpp", [0.5, 1.0, 2.0])
def test_encoder_lossy_rate_and_openjpeg_agreement(bpp):
    a = synth.image(160, 192, 3, 9)
    cs = synth.encode(a, levels=5, cb=32, lossy=True, bpp=bpp)
    assert len(cs) * 8 / (160 * 192) <= bpp * 1.02
    d = oracle.decode(cs).transpose(1, 2, 0)
    mse = np.mean((d - a.astype(np.int32)) ** 2)
    assert 10 * np.log10(255 ** 2 / mse) > 18  # noisy content: low PSNR at 0.5 bpp
    if imaging.pillow_j2k() is not None:
        r = imaging.decode_openjpeg(cs).astype(np.int32)
        assert np.abs(r - d).max() <= 1


def _random_block(rng, w, h, bits):
    a = rng.integers(-(1 << bits) + 1, 1 << bits, (h, w)).astype(np.int32)
    a[rng.uniform(size=(h, w)) < rng.uniform(0, 0.9)] = 0  # sparse regions exercise run-length mode
    return a


@pytest.mark.parametrize("seed", range(4))
def test_tier1_round_trip_all_block_shapes(seed):
    import ctypes
    rng = np.random.default_rng(seed)
    L = synth._enc_lib()
    for _ in range(40):
        w, h = int(rng.integers(1, 65)), int(rng.integers(1, 65))
        orient = int(rng.integers(0, 4))
        a = _random_block(rng, w, h, int(rng.integers(1, 16)))
        buf = np.zeros(1 << 16, np.uint8)
        n = ctypes.c_int(0)
        nb = L.tce_t1_encode_block(a.ctypes.data, w, h, orient, buf.ctypes.data, buf.size, ctypes.byref(n))
        assert nb >= 0
        if nb == 0:
            assert not a.any()
            continue
        mag, sign, low, _ = oracle.t1_decode(bytes(buf[:n.value]), w, h, orient, nb, 3 * nb - 2)
        got = np.where(sign.astype(bool), -mag, mag)
        assert np.array_equal(got, a), (w, h, orient)


@pytest.mark.parametrize("seed", range(3))
def test_tier1_truncated_passes_decode_the_prefix(seed):
    import ctypes
    rng = np.random.default_rng(10 + seed)
    L = synth._enc_lib()
    for _ in range(25):
        w, h = int(rng.integers(4, 65)), int(rng.integers(4, 65))
        a = _random_block(rng, w, h, 10)
        buf = np.zeros(1 << 16, np.uint8)
        nb = ctypes.c_int(0)
        full = None
        for keep in (1, 2, 3, 4, 7):
            n = L.tce_t1_encode_truncated(a.ctypes.data, w, h, 0, keep, buf.ctypes.data, buf.size, ctypes.byref(nb))
            if nb.value == 0:
                break
            k = min(keep, 3 * nb.value - 2)
            mag, sign, low, _ = oracle.t1_decode(bytes(buf[:n]), w, h, 0, nb.value, k)
            # every revealed bit-plane matches the source magnitude above the last coded plane
            p = nb.value - 1 - (k + 1) // 3
            src = np.abs(a)
            assert np.array_equal(mag >> max(p + 1, 0), src >> max(p + 1, 0)), (w, h, keep)
