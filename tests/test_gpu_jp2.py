"""JP2 files (tests/golden/jp2_*, written by OpenJPEG) through the GPU decode
paths: the fused per-image path and a batch mixing them with raw codestreams
equal the source (lossless) or OpenJPEG's decode within 1 (lossy); a file the
JP2 layer rejects fails the decode without launching anything."""
import numpy as np
import pytest

import thousandthchicken_amd as tca
from tests.test_jp2 import NAMES, edit_header, load, raw_codestream

pytestmark = pytest.mark.gpu


def _check(got, name, z):
    ref = z["source"] if name.endswith("_ll") else z["openjpeg"]
    d = np.abs(got.reshape(ref.shape).astype(np.int32) - ref).max()
    assert d <= (0 if name.endswith("_ll") else 1), (name, d)


@pytest.mark.parametrize("name", NAMES)
def test_fused_decode_of_jp2(gpu_decoder, name):
    data, z = load(name)
    print(f"jp2 {name}", flush=True)
    _check(gpu_decoder.decode(data), name, z)
    assert np.array_equal(gpu_decoder.decode(data), gpu_decoder.decode(raw_codestream(data)))


def test_batch_of_jp2_and_raw_codestreams():
    cases = [load(n) + (n,) for n in NAMES]
    streams = [d for d, _, _ in cases] + [raw_codestream(d) for d, _, _ in cases]
    dec = tca.Decoder()
    try:
        dec.prepare(streams)
        dec.run()
        dec.sync()
        for i, (data, z, name) in enumerate(cases):
            _check(dec.fetch(i), name, z)
            assert np.array_equal(dec.fetch(i), dec.fetch(i + len(cases)))
    finally:
        dec.close()


def test_rejected_jp2_fails_the_decode(gpu_decoder):
    data, _ = load("rgb_ll")
    bad = edit_header(data, lambda sub: sub + [(b"cmap", bytes([0, 0, 1, 0]))])
    with pytest.raises(tca.TCError):
        gpu_decoder.decode(bad)
    dec = tca.Decoder()
    try:
        with pytest.raises(tca.TCError):
            dec.prepare([data, bad])
    finally:
        dec.close()
