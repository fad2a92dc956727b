"""Resolution-reduced decode on the GPU against the oracle's reduced decode
(itself pinned by the low-pass closed form in test_reduce.py): lossless
bit-exact, lossy within 1 LSB, per image, for multi-tile images, and in a
batch."""
import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle
from thousandthchicken_amd import synth

pytestmark = pytest.mark.gpu


def _check(dec, cs, k, lim):
    dec.set_reduce(k)
    got = dec.decode(cs).astype(np.int32)
    ref = oracle.decode(cs, reduce=k).transpose(1, 2, 0)
    assert got.shape == ref.shape, (k, got.shape, ref.shape)
    d = int(np.abs(got - ref).max()) if got.size else 0
    assert d <= lim, (k, d)


@pytest.mark.parametrize("case", ["rgb_lossless", "gray16_lossless", "rgb_lossy", "tiled_lossless"])
def test_reduced_per_image(case):
    print(f"reduce {case}", flush=True)
    if case == "rgb_lossless":
        cs, levels, lim = synth.encode(synth.image(301, 257, 3, 8), levels=5, cb=64), 5, 0
    elif case == "gray16_lossless":
        cs, levels, lim = synth.encode(synth.image(140, 333, 1, 9, 16), bits=16, levels=6, cb=32), 6, 0
    elif case == "rgb_lossy":
        cs, levels, lim = synth.encode(synth.image(256, 300, 3, 10), levels=5, cb=32, lossy=True, bpp=1.0), 5, 1
    else:
        cs, levels, lim = synth.encode(synth.image(200, 330, 3, 11), levels=4, cb=32, tile=(128, 96)), 4, 0
    dec = tca.Decoder()
    for k in range(levels + 1):
        _check(dec, cs, k, lim)
    dec.close()


def test_reduced_batch_equals_single():
    streams = [synth.encode(synth.image(90 + 17 * i, 120 + 9 * i, 3 if i % 2 else 1, 20 + i), levels=4, cb=32,
                            lossy=bool(i % 3 == 1), bpp=1.5) for i in range(6)]
    one = tca.Decoder()
    one.set_reduce(2)
    single = [one.decode(cs) for cs in streams]
    one.close()
    dec = tca.Decoder()
    dec.set_reduce(2)
    dec.prepare(streams)
    dec.run()
    dec.sync()
    for i in range(len(streams)):
        assert np.array_equal(dec.fetch(i), single[i]), i
    dec.close()


def test_reduce_beyond_levels_fails():
    cs = synth.encode(synth.image(64, 64, 1, 3), levels=3, cb=32)
    dec = tca.Decoder()
    dec.set_reduce(4)
    with pytest.raises(tca.TCError):
        dec.prepare([cs])
    dec.close()
