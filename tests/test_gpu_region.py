"""Region decode on the GPU: the output window equals the same window of the
full decode -- bit for bit, lossless and lossy, since every sample inside is
computed from the same coefficients by the same arithmetic -- and the oracle's
full decode cropped (lossless exact, lossy within 1).  Windows cover single
samples, image edges, tile boundaries, odd image/tile origins, the whole image,
resolution-reduced images and batches."""
import os

import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle
from thousandthchicken_amd import synth

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _streams():
    out = []
    src = synth.image(300, 420, 3, 71)
    out.append(("ll_tiles", synth.encode(src, levels=5, cb=32, tile=(128, 160)), False))
    out.append(("ly_one_tile", synth.encode(src, levels=5, cb=64, lossy=True, bpp=1.5), True))
    g = synth.image(257, 190, 1, 72, 12)
    out.append(("ll_gray12", synth.encode(g, bits=12, levels=4, cb=32), False))
    for name in ("openjpeg_ll_tiles_offset.j2k", "openjpeg_ly_rgb_layers.j2k", "openjpeg_ll_prec_pcrl_tiles.j2k"):
        out.append((name, open(os.path.join(GOLDEN, name), "rb").read(), "_ly_" in name))
    return out


STREAMS = _streams()


def _windows(w, h, seed):
    rng = np.random.default_rng(seed)
    win = [(0, 0, 1, 1), (w - 1, h - 1, 1, 1), (0, 0, w, h), (w // 3, h // 4, 17, 9), (0, h // 2, w, 3),
           (w // 2, 0, 5, h), (w - 40, h - 30, 100, 100)]
    for _ in range(6):
        x, y = int(rng.integers(0, w)), int(rng.integers(0, h))
        win.append((x, y, int(rng.integers(1, w - x + 1)), int(rng.integers(1, h - y + 1))))
    return win


@pytest.mark.parametrize("reduce", [0, 1])
@pytest.mark.parametrize("idx", range(len(STREAMS)), ids=[s[0] for s in STREAMS])
def test_region_equals_full_decode_window(gpu_decoder, idx, reduce):
    name, cs, lossy = STREAMS[idx]
    full_dec = tca.Decoder()
    full_dec.set_reduce(reduce)
    full = full_dec.decode(cs)
    full_dec.close()
    ref = oracle.decode(cs, reduce=reduce).transpose(1, 2, 0).astype(np.int32)
    h, w = full.shape[:2]
    dec = tca.Decoder()
    dec.set_reduce(reduce)
    try:
        for (x, y, ww, hh) in _windows(w, h, idx * 10 + reduce):
            print(f"region {name} reduce={reduce} ({x},{y},{ww},{hh})", flush=True)
            dec.set_region(x, y, ww, hh)
            out = dec.decode(cs)
            exp = full[y:y + hh, x:x + ww]
            assert out.shape == exp.shape, (out.shape, exp.shape)
            assert np.array_equal(out, exp), (x, y, ww, hh)
            d = np.abs(out.astype(np.int32) - ref[y:y + hh, x:x + ww])
            assert d.max() <= (1 if lossy else 0)
    finally:
        dec.close()


def test_region_in_a_batch():
    cases = STREAMS[:4]
    fulls = []
    for _, cs, _ in cases:
        d = tca.Decoder()
        fulls.append(d.decode(cs))
        d.close()
    dec = tca.Decoder()
    try:
        dec.set_region(50, 40, 120, 90)
        dec.prepare([c[1] for c in cases])
        dec.run()
        dec.sync()
        for i, f in enumerate(fulls):
            assert np.array_equal(dec.fetch(i), f[40:130, 50:170])
    finally:
        dec.close()


def test_region_outside_the_image_fails():
    cs = STREAMS[0][1]
    dec = tca.Decoder()
    try:
        dec.set_region(10000, 0, 5, 5)
        with pytest.raises(tca.TCError):
            dec.prepare([cs])
    finally:
        dec.close()
