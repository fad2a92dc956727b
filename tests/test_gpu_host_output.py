"""Zero-copy host output: with set_host_output the final-level kernel writes
the images into pinned host memory directly; they equal the device-buffer
output, per image, batched (the buffer grows with the batch), through the
in-place view and through a captured graph replayed several times."""
import os

import numpy as np
import pytest

import thousandthchicken_amd as tca
from thousandthchicken_amd import synth

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _streams():
    # sub-sampled streams decode to per-component planes (tests/test_gpu_openjpeg_pins.py)
    names = sorted(f for f in os.listdir(GOLDEN) if f.endswith(".j2k") and not f.startswith("pkg_sub_"))
    out = [open(os.path.join(GOLDEN, n), "rb").read() for n in names]
    out.append(synth.encode(synth.image(300, 260, 3, 5), levels=4, cb=32, lossy=True, bpp=1.0))
    return out


def test_host_output_equals_device_output(gpu_decoder):
    streams = _streams()
    dec = tca.Decoder()
    try:
        dec.set_host_output(True)
        for cs in streams:
            assert np.array_equal(dec.decode(cs), gpu_decoder.decode(cs))
        # batched, larger than any single image: the pinned buffer grows
        dec.prepare(streams)
        dec.run()
        dec.sync()
        for i, cs in enumerate(streams):
            ref = gpu_decoder.decode(cs)
            assert np.array_equal(dec.fetch(i), ref)
            assert np.array_equal(dec.output_view(i), ref)
    finally:
        dec.close()


def test_host_output_through_graph_replays(gpu_decoder):
    streams = _streams()[:6]
    refs = [gpu_decoder.decode(cs) for cs in streams]
    dec = tca.Decoder()
    try:
        dec.set_host_output(True)
        dec.prepare(streams)
        dec.capture_graph()
        for rep in range(3):
            dec.launch_graph()
            dec.sync()
            for i, r in enumerate(refs):
                assert np.array_equal(dec.output_view(i), r), (rep, i)
            # eager work between replays
            gpu_decoder.decode(streams[rep])
    finally:
        dec.close()
