"""tc_decode_noalloc (the per-image decode that allocates nothing, its working
memory in one static device array per GPU) called through the C ABI: its
output equals tc_decoder_decode's on two calls in a row, the second with a
smaller image, and an image too large for the array fails cleanly without
disturbing the next call."""
import ctypes

import numpy as np
import pytest

import thousandthchicken_amd as tca
from thousandthchicken_amd import synth

pytestmark = pytest.mark.gpu


def _noalloc(stream, cs, nbytes):
    dev = tca.default_device()
    buf = tca.DeviceBuffer(dev, max(nbytes, 256))
    try:
        info = tca.ImageInfo()
        tca.check(tca.lib().tc_decode_noalloc(stream, cs, len(cs), buf.ptr, nbytes, ctypes.byref(info)),
                  "tc_decode_noalloc")
        return buf.download(stream, np.uint8, nbytes), info
    finally:
        buf.free()


def _existing(cs, nbytes):
    dec = tca.Decoder()
    try:
        out = np.empty(nbytes, np.uint8)
        tca.check(tca.lib().tc_decoder_decode(dec.handle, cs, len(cs), out.ctypes.data, nbytes, None), "decode")
        return out
    finally:
        dec.close()


def _nbytes(cs):
    i = tca.probe(cs)
    return i.width * i.height * i.ncomp * i.bytes_per_sample


def test_noalloc_equals_decoder_decode_twice_then_smaller():
    L = tca.lib()
    stream = ctypes.c_void_p()
    tca.check(L.tc_stream_create(tca.default_device().handle, ctypes.byref(stream)), "stream")
    try:
        cases = [
            # larger first: a lossy multi-block 9/7 image, then a smaller lossless one
            synth.encode(synth.image(600, 520, 3, 71), levels=5, cb=64, lossy=True, bpp=1.5),
            synth.encode(synth.image(96, 80, 3, 72), levels=4, cb=32),
            # 16-bit single component, 64x64 lossless
            synth.encode(synth.image(300, 260, 1, 73, 16)[..., 0], bits=16, levels=5, cb=64),
            # styled code-blocks decode under lockstep packing, whose tier-1 state is in the array too
            synth.encode(synth.image(200, 150, 3, 75), levels=4, cb=32, cbsty=63),
        ]
        for k, cs in enumerate(cases):
            n = _nbytes(cs)
            print(f"noalloc case {k}: {len(cs)} bytes, output {n}", flush=True)
            got, info = _noalloc(stream, cs, n)
            ref = _existing(cs, n)
            assert np.array_equal(got, ref), k
            assert info.width * info.height * info.ncomp * info.bytes_per_sample == n
    finally:
        L.tc_stream_destroy(stream)


def test_noalloc_rejects_images_beyond_its_array():
    L = tca.lib()
    stream = ctypes.c_void_p()
    tca.check(L.tc_stream_create(tca.default_device().handle, ctypes.byref(stream)), "stream")
    try:
        # 8448 x 8192 samples of 4 bytes each: more than the 256 MiB array for the planes alone
        big = synth.encode(np.zeros((8192, 8448), np.uint8), levels=2, cb=64)
        buf = tca.DeviceBuffer(tca.default_device(), 256)
        try:
            st = L.tc_decode_noalloc(stream, big, len(big), buf.ptr, 256, None)
        finally:
            buf.free()
        assert st == 8, st  # TC_ERR_CAPACITY
        assert b"static array" in L.tc_last_error()
        small = synth.encode(synth.image(64, 48, 3, 74), levels=3, cb=32)
        got, _ = _noalloc(stream, small, _nbytes(small))
        assert np.array_equal(got, _existing(small, _nbytes(small)))
    finally:
        L.tc_stream_destroy(stream)
