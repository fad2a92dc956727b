"""Sub-sampled components on the GPU: each component decodes at its own size
into its own output plane, equal to the oracle (lossless: also the encoder's
source samples img[::YRsiz, ::XRsiz, c]; lossy: within 1), per image, at a
reduced resolution, and batched beside interleaved images."""
import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle
from tests.test_subsampling import CASES
from thousandthchicken_amd import synth

pytestmark = pytest.mark.gpu


def _stream(case, lossy=False):
    name, shape, sub, tile, bits = case
    src = synth.image(shape[0], shape[1], shape[2], 91, bits)
    cs = synth.encode(src, bits=bits, levels=3, cb=32, subsampling=sub, tile=tile, lossy=lossy, bpp=3.0)
    return src, cs


@pytest.mark.parametrize("lossy", [False, True])
@pytest.mark.parametrize("idx", range(len(CASES)), ids=[c[0] for c in CASES])
def test_subsampled_decode_equals_oracle(gpu_decoder, idx, lossy):
    name, shape, sub, tile, bits = CASES[idx]
    src, cs = _stream(CASES[idx], lossy)
    print(f"subsampled {name} lossy={lossy}", flush=True)
    got = gpu_decoder.decode(cs)
    ref = oracle.decode(cs)
    assert isinstance(got, list) and len(got) == len(sub)
    for c, (dx, dy) in enumerate(sub):
        d = np.abs(got[c].astype(np.int32) - ref[c])
        assert d.max() <= (1 if lossy else 0), c
        if not lossy:
            assert np.array_equal(got[c], src[::dy, ::dx, c])


@pytest.mark.parametrize("reduce", [1, 2])
def test_subsampled_reduced_decode_equals_oracle(reduce):
    src, cs = _stream(CASES[3])
    dec = tca.Decoder()
    try:
        dec.set_reduce(reduce)
        got = dec.decode(cs)
    finally:
        dec.close()
    ref = oracle.decode(cs, reduce=reduce)
    for c in range(len(ref)):
        assert np.array_equal(got[c], ref[c]), c


def test_subsampled_and_interleaved_in_one_batch():
    streams = [_stream(c)[1] for c in CASES]
    plain = synth.image(80, 90, 3, 92)
    streams.append(synth.encode(plain, levels=3, cb=32))
    dec = tca.Decoder()
    try:
        dec.prepare(streams)
        dec.run()
        dec.sync()
        for i, cs in enumerate(streams):
            got, ref = dec.fetch(i), oracle.decode(cs)
            if isinstance(ref, list):
                for c in range(len(ref)):
                    assert np.array_equal(got[c], ref[c]), (i, c)
            else:
                assert np.array_equal(got, ref.transpose(1, 2, 0)), i
    finally:
        dec.close()


def test_region_and_staged_decode_of_subsampled_images_fail_cleanly():
    _, cs = _stream(CASES[0])
    dec = tca.Decoder()
    try:
        dec.set_region(0, 0, 10, 10)
        with pytest.raises(tca.TCError):
            dec.prepare([cs])
    finally:
        dec.close()
    with pytest.raises(tca.TCError):
        tca.decode_staged(cs)
