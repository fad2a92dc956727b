"""Decodes of the package encoder's streams: 12-bit multi-tile lossless
(the cfg3 layout at reduced size), 16-bit single component, and a reduced
mixed batch (the cfg4 distribution scaled down), against the source / oracle."""
import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle
from thousandthchicken_amd import synth, workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,bits,kw", [
    ((540, 960, 3), 12, dict(levels=6, cb=32, tile=(256, 256))),
    ((300, 520, 1), 16, dict(levels=6, cb=64)),
    ((257, 333, 3), 10, dict(levels=5, cb=32, tile=(128, 64))),
])
def test_high_bit_depth_lossless_bit_exact(gpu_decoder, shape, bits, kw):
    h, w, c = shape
    a = synth.image(h, w, c, 31, bits)
    cs = synth.encode(a, bits=bits, **kw)
    print(f"case {shape} {bits}-bit {len(cs)} bytes", flush=True)
    out = gpu_decoder.decode(cs)
    assert out.dtype == np.uint16
    assert np.array_equal(out.reshape(a.shape), a)


def test_mixed_batch_reduced(gpu_decoder):
    streams = workloads.cfg4(n=24, scale=0.15)
    lay = workloads.cfg4_layout(24)
    dec = tca.Decoder()
    dec.prepare(streams)
    dec.run()
    dec.sync()
    for i, cs in enumerate(streams):
        out = dec.fetch(i).astype(np.int32)
        ref = oracle.decode(cs).transpose(1, 2, 0)
        lim = 0 if lay[i]["lossless"] else 1
        assert np.abs(out - ref).max() <= lim, i
        if lay[i]["lossless"]:
            src = workloads.source_image("cfg4", i, scale=0.15)
            assert np.array_equal(out.reshape(src.shape), src.astype(np.int32))
    dec.close()
