"""GPU forward transform (encoder front end) against the oracle.

- The DWTForward53/97 drop-in classes on one plane equal the oracle's forward
  DWT (T.800 F.4): 5/3 bit-exact, 9/7 within FP32 rounding of the FP64 oracle,
  over odd sizes, length-1 axes and every level count.
- Forward then inverse through the drop-in classes returns the plane (5/3
  exactly), tying the forward kernels to the inverse ones the decoder fixtures
  pin.
- The fused image path (DC shift + forward RCT/ICT + first level, then the
  other levels) equals the oracle's forward DWT of the colour-transformed
  image, and inverting it (GPU inverse DWT, inverse RCT, shift back) gives the
  image again.
"""
import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle
from thousandthchicken_amd import synth

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (1, 7), (7, 1), (2, 3), (5, 9), (64, 64), (65, 129), (257, 300), (31, 512)]


def _levels(h, w):
    return sorted({0, 1, min(3, max(1, max(h, w).bit_length())), 6})


def _tol97(ref):
    return 2e-4 * max(1.0, float(np.abs(ref).max()))


@pytest.mark.parametrize("hw", SIZES, ids=[f"{h}x{w}" for h, w in SIZES])
def test_drop_in_forward53_equals_oracle(hw):
    h, w = hw
    rng = np.random.default_rng(h * 1000 + w)
    a = rng.integers(-2048, 2048, (h, w)).astype(np.int32)
    for lv in _levels(h, w):
        print(f"fdwt53 {h}x{w} levels={lv}", flush=True)
        got = tca.dwt_plane(a, lv, forward=True)
        assert np.array_equal(got, oracle.fdwt53(a, 0, 0, lv)), lv
        assert np.array_equal(tca.dwt_plane(got, lv, forward=False), a), lv


@pytest.mark.parametrize("hw", SIZES, ids=[f"{h}x{w}" for h, w in SIZES])
def test_drop_in_forward97_within_fp32_of_oracle(hw):
    h, w = hw
    rng = np.random.default_rng(h * 1000 + w + 7)
    a = rng.uniform(-2048, 2048, (h, w)).astype(np.float32)
    for lv in _levels(h, w):
        print(f"fdwt97 {h}x{w} levels={lv}", flush=True)
        got = tca.dwt_plane(a, lv, forward=True)
        ref = oracle.fdwt97(a.astype(np.float64), 0, 0, lv)
        assert np.abs(got - ref).max() <= _tol97(ref), lv
        back = tca.dwt_plane(got, lv, forward=False)
        assert np.abs(back - a).max() <= _tol97(a), lv


def _rct(x):
    r, g, b = (x[..., i].astype(np.int64) for i in range(3))
    return np.stack([(r + 2 * g + b) >> 2, b - g, r - g])


def _irct(y):
    g = y[0] - ((y[1] + y[2]) >> 2)
    return np.stack([y[2] + g, g, y[1] + g], -1)


def _ict(x):
    r, g, b = (x[..., i].astype(np.float64) for i in range(3))
    return np.stack([0.299 * r + 0.587 * g + 0.114 * b,
                     -0.16875 * r - 0.33126 * g + 0.5 * b,
                     0.5 * r - 0.41869 * g - 0.08131 * b])


CASES = [
    ("rgb8_l5", (200, 300, 3), 8, 5),
    ("rgb8_odd_l3", (97, 131, 3), 8, 3),
    ("rgb8_l0", (40, 33, 3), 8, 0),
    ("gray12_l4", (150, 90, 1), 12, 4),
    ("gray16_l6", (64, 257, 1), 16, 6),
    ("rgb10_l1", (33, 65, 3), 10, 1),
]


@pytest.mark.parametrize("name,shape,bits,levels", CASES, ids=[c[0] for c in CASES])
def test_forward_transform_53_equals_oracle_and_inverts(name, shape, bits, levels):
    h, w, c = shape
    img = synth.image(h, w, c, 61, bits)
    print(f"forward 5/3 {name}", flush=True)
    got = tca.forward_transform(img, levels=levels, irreversible=False, bits=bits)
    x = img.astype(np.int64) - (1 << (bits - 1))
    comps = _rct(x) if c == 3 else np.moveaxis(x, -1, 0)
    for k in range(c):
        assert np.array_equal(got[k], oracle.fdwt53(comps[k].astype(np.int32), 0, 0, levels)), k
    # back through the inverse drop-in classes and the inverse RCT
    back = np.stack([tca.dwt_plane(got[k], levels, forward=False) for k in range(c)]).astype(np.int64)
    pix = (_irct(back) if c == 3 else np.moveaxis(back, 0, -1)) + (1 << (bits - 1))
    assert np.array_equal(pix, img.astype(np.int64))


@pytest.mark.parametrize("name,shape,bits,levels", CASES, ids=[c[0] for c in CASES])
def test_forward_transform_97_within_fp32_of_oracle(name, shape, bits, levels):
    h, w, c = shape
    img = synth.image(h, w, c, 62, bits)
    print(f"forward 9/7 {name}", flush=True)
    got = tca.forward_transform(img, levels=levels, irreversible=True, bits=bits)
    x = img.astype(np.float64) - (1 << (bits - 1))
    comps = _ict(x) if c == 3 else np.moveaxis(x, -1, 0)
    for k in range(c):
        ref = oracle.fdwt97(comps[k], 0, 0, levels)
        assert np.abs(got[k] - ref).max() <= _tol97(ref), k


@pytest.mark.parametrize("shape", [0, 1, 2, 3])
def test_tile_shapes_agree(shape):
    img = synth.image(130, 170, 3, 63)
    ref = tca.forward_transform(img, levels=4)
    f = tca.Forward()
    try:
        f.set_shape(shape)
        f.prepare(170, 130, 3, 8, 4, False, True)
        f.upload(img)
        f.run()
        f.sync()
        assert np.array_equal(f.download(), ref)
    finally:
        f.close()
