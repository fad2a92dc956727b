"""Precincts (T.800 B.6) and the five progression orders (B.12.1) with more
than one precinct per resolution, on codestreams written by OpenJPEG (Pillow):
the oracle decodes the lossless ones to their source, and the product's parser
lists the same code-blocks (passes, zero bit-planes, lengths and a hash of the
codeword bytes) as the oracle.  The two order the position-driven progressions
differently: the oracle steps through reference-grid positions as B.12.1.3-5
write them, the parser sorts precincts by the position where each starts."""
import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle
from tests import imaging
from thousandthchicken_amd import synth

pytestmark = pytest.mark.skipif(imaging.pillow_j2k() is None, reason="OpenJPEG (Pillow) not available")

ORDERS = ["LRCP", "RLCP", "RPCL", "PCRL", "CPRL"]
LAYOUTS = [
    ("p64_cb16", dict(num_resolutions=5, codeblock_size=(16, 16), precinct_size=(64, 64))),
    ("p32_cb16_tiles_off", dict(num_resolutions=4, codeblock_size=(16, 16), precinct_size=(32, 32),
                                tile_size=(96, 80), offset=(7, 2), tile_offset=(1, 1))),
    ("p128_cb32_tiles", dict(num_resolutions=6, codeblock_size=(32, 32), precinct_size=(128, 128),
                             tile_size=(128, 128), offset=(3, 5))),
    ("p64_cb32_layers", dict(num_resolutions=4, codeblock_size=(32, 32), precinct_size=(64, 64),
                             quality_layers=[30, 10, 0])),
]


def _key(c):
    return (c.tile, c.comp, c.res, c.band, c.x0, c.y0, c.x1, c.y1, c.npasses, c.zero_bp, c.mb, c.nbytes, c.crc)


@pytest.mark.parametrize("layout", LAYOUTS, ids=[x[0] for x in LAYOUTS])
@pytest.mark.parametrize("order", ORDERS)
def test_precinct_streams_decode_and_parse_alike(order, layout):
    src = synth.image(300, 200, 3, 6)
    cs = imaging.encode_openjpeg(src, progression=order, mct=1, **layout[1])
    # more than one precinct in the finest resolutions: the COD carries them
    i = cs.index(b"\xff\x52")
    assert cs[i + 4] & 1
    assert np.array_equal(oracle.decode(cs).transpose(1, 2, 0), src)
    assert [_key(c) for c in tca.list_cblks(cs)] == [_key(c) for c in oracle.list_cblks(cs)]


@pytest.mark.parametrize("order", ORDERS)
def test_lossy_precinct_streams_within_one_of_openjpeg(order):
    src = synth.image(180, 230, 3, 9)
    cs = imaging.encode_openjpeg(src, progression=order, mct=1, irreversible=True, quality_mode="rates",
                                 quality_layers=[20, 6], num_resolutions=5, codeblock_size=(32, 32),
                                 precinct_size=(128, 128))
    ref = imaging.decode_openjpeg(cs).astype(np.int32)
    got = oracle.decode(cs).transpose(1, 2, 0)
    assert np.abs(got - ref).max() <= 1
    assert [_key(c) for c in tca.list_cblks(cs)] == [_key(c) for c in oracle.list_cblks(cs)]


def test_precinct_exponent_zero_above_lowest_resolution_is_rejected():
    src = synth.image(64, 64, 1, 3)[..., 0]
    cs = bytearray(imaging.encode_openjpeg(src, num_resolutions=3, codeblock_size=(16, 16),
                                           precinct_size=(64, 64)))
    i = cs.index(b"\xff\x52")
    # COD: marker, Lcod, Scod, SGcod (4), levels, xcb, ycb, style, transform, then PPx|PPy per resolution
    pp = i + 2 + 2 + 1 + 4 + 5
    cs[pp + 1] = 0x00
    with pytest.raises(tca.TCError):
        tca.probe(bytes(cs))
    with pytest.raises(oracle.OracleError):
        oracle.decode(bytes(cs))
