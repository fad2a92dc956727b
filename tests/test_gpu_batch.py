"""Batched decode: many images of uneven size in one launch per stage give, per
image, exactly what the per-image path gives (and the oracle agrees)."""
import os

import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _streams():
    # sub-sampled streams decode to per-component planes (tests/test_gpu_openjpeg_pins.py)
    names = sorted(f for f in os.listdir(GOLDEN) if f.endswith(".j2k") and not f.startswith("pkg_sub_"))
    return [open(os.path.join(GOLDEN, n), "rb").read() for n in names]


@pytest.mark.parametrize("bpw", [0, 1, -64, -32, -16])
def test_batch_equals_single(gpu_decoder, bpw):
    streams = _streams()
    print(f"batch of {len(streams)} bpw={bpw}", flush=True)
    singles = [gpu_decoder.decode(cs) for cs in streams]
    dec = tca.Decoder()
    dec.set_blocks_per_wave(True, bpw)
    dec.set_blocks_per_wave(False, bpw)
    dec.prepare(streams)
    dec.run()
    dec.sync()
    for i, cs in enumerate(streams):
        out = dec.fetch(i)
        ref = oracle.decode(cs).transpose(1, 2, 0)
        assert np.abs(out.astype(np.int32) - ref).max() <= (1 if tca.probe(cs).reversible == 0 else 0)
        if tca.probe(cs).reversible:
            assert np.array_equal(out, singles[i])
    dec.close()


def test_batch_graph_replay_matches_eager(gpu_decoder):
    streams = _streams()
    dec = tca.Decoder()
    dec.prepare(streams)
    dec.run()
    dec.sync()
    eager = [dec.fetch(i) for i in range(len(streams))]
    dec.capture_graph()
    for rep in range(3):
        dec.launch_graph()
        dec.sync()
        for i in range(len(streams)):
            assert np.array_equal(dec.fetch(i), eager[i]), (rep, i)
        dec.run()  # eager work between replays
        dec.sync()
    dec.close()


def _shape_count():
    import ctypes
    n = ctypes.c_int(0)
    tca.check(tca.lib().tc_idwt_tile_shape(0, None, None, ctypes.byref(n)), "shapes")
    return n.value


@pytest.mark.parametrize("shape", range(4))
def test_dwt_tile_shapes(gpu_decoder, shape):
    assert _shape_count() == 4
    streams = _streams()
    print(f"dwt tile shape {shape}", flush=True)
    dec = tca.Decoder()
    t = dec.tuning()
    t.dwt_shape = shape
    dec.set_tuning(t)
    dec.prepare(streams)
    dec.run()
    dec.sync()
    for i, cs in enumerate(streams):
        ref = oracle.decode(cs).transpose(1, 2, 0)
        lim = 0 if tca.probe(cs).reversible else 1
        assert np.abs(dec.fetch(i).astype(np.int32) - ref).max() <= lim, (shape, i)
    dec.close()


def test_autotune_keeps_outputs(gpu_decoder):
    streams = _streams()
    dec = tca.Decoder()
    dec.prepare(streams)
    dec.run()
    dec.sync()
    before = [dec.fetch(i) for i in range(len(streams))]
    best = dec.autotune(reps=1)
    print("autotune:", best.as_dict(), flush=True)
    packings = (1, tca.BPW_LOCKSTEP, tca.BPW_LOCKSTEP32, tca.BPW_LOCKSTEP16)
    assert best.bpw_small in packings and best.bpw_large in packings
    assert 0 <= best.dwt_shape < 4 and best.t1_ms > 0
    dec.run()
    dec.sync()
    for i, cs in enumerate(streams):
        lim = 0 if tca.probe(cs).reversible else 1
        assert np.abs(dec.fetch(i).astype(np.int32) - before[i].astype(np.int32)).max() <= lim
    dec.close()
