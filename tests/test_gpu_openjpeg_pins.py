"""The GPU decode paths against OpenJPEG 2.5.4 on the pkg_* fixtures: code-block
styles (BYPASS raw segments included), PPT/PPM packet headers and sub-sampled
components.  The fused per-image path, the staged stage-class path and the
batched path (every fixture in one batch, under several code-block packings)
must each equal OpenJPEG's decode exactly for lossless streams and within 1
for lossy ones."""
import numpy as np
import pytest

import thousandthchicken_amd as tca
from tests import pkg_fixtures as pf

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", pf.names())
def test_fused_decode_equals_openjpeg(gpu_decoder, name):
    cs, ref, meta = pf.load(name)
    print(f"pkg fused {name}", flush=True)
    d = pf.max_diff(gpu_decoder.decode(cs), ref, meta)
    assert d <= pf.tolerance(meta), (name, d)


@pytest.mark.parametrize("name", [n for n in pf.names() if not n.startswith("sub_")])
def test_staged_decode_equals_openjpeg(name):
    cs, ref, meta = pf.load(name)
    print(f"pkg staged {name}", flush=True)
    d = pf.max_diff(tca.decode_staged(cs), ref, meta)
    assert d <= pf.tolerance(meta), (name, d)


@pytest.mark.parametrize("bpw", [0, 1, tca.BPW_LOCKSTEP16])
def test_batched_decode_equals_openjpeg(bpw):
    cases = [pf.load(n) for n in pf.names()]
    print(f"pkg batch of {len(cases)} bpw={bpw}", flush=True)
    dec = tca.Decoder()
    try:
        dec.set_blocks_per_wave(True, bpw)
        dec.set_blocks_per_wave(False, bpw)
        dec.prepare([c[0] for c in cases])
        dec.run()
        dec.sync()
        for i, (cs, ref, meta) in enumerate(cases):
            d = pf.max_diff(dec.fetch(i), ref, meta)
            assert d <= pf.tolerance(meta), (i, d)
    finally:
        dec.close()
