"""A prepare that fails leaves nothing to run: the decoder reports an error for
run, graph capture, fetch and output queries instead of launching a plan that
no longer matches its buffers, and the next good prepare decodes normally."""
import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle
from thousandthchicken_amd import synth

pytestmark = pytest.mark.gpu


def _good():
    src = synth.image(64, 80, 3, 61)
    return src, synth.encode(src, levels=3, cb=32)


def test_failed_plan_is_not_run():
    src, good = _good()
    # 128x32 code-blocks are valid T.800 (area 4096) but the planner takes at most 64x64
    bad = synth.encode(synth.image(70, 150, 3, 62), levels=3, cb=(128, 32))
    assert np.array_equal(oracle.decode(bad).transpose(1, 2, 0), synth.image(70, 150, 3, 62))
    dec = tca.Decoder()
    try:
        dec.prepare([good])
        dec.run()
        dec.sync()
        assert np.array_equal(dec.fetch(0), src)
        print("prepare of a 128x32-block stream", flush=True)
        with pytest.raises(tca.TCError):
            dec.prepare([good, bad])
        for call in (dec.run, dec.capture_graph, dec.launch_graph, lambda: dec.fetch(0)):
            with pytest.raises(tca.TCError):
                call()
        assert dec.stats().images == 0
        dec.prepare([good])
        dec.run()
        dec.sync()
        assert np.array_equal(dec.fetch(0), src)
    finally:
        dec.close()


def test_failed_parse_drops_the_previous_plan():
    src, good = _good()
    dec = tca.Decoder()
    try:
        dec.prepare([good])
        with pytest.raises(tca.TCError):
            dec.prepare([good[: len(good) // 3]])
        with pytest.raises(tca.TCError):
            dec.run()
        dec.prepare([good])
        dec.run()
        dec.sync()
        assert np.array_equal(dec.fetch(0), src)
    finally:
        dec.close()
