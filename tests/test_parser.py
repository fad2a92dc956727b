"""The product's tier-2 parser (C++ in libtc.so, run on the host) against the
oracle's tables, plus its error handling on damaged codestreams."""
import os

import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(f for f in os.listdir(GOLDEN) if f.endswith(".j2k"))


def _key(c):
    return (c.tile, c.comp, c.res, c.band, c.x0, c.y0, c.x1, c.y1, c.npasses, c.zero_bp, c.mb, c.nbytes, c.crc)


@pytest.mark.parametrize("name", FIXTURES)
def test_codeblock_table_equals_oracle(name):
    cs = open(os.path.join(GOLDEN, name), "rb").read()
    a = [_key(c) for c in oracle.list_cblks(cs)]
    b = [_key(c) for c in tca.list_cblks(cs)]
    assert a == b


@pytest.mark.parametrize("name", FIXTURES)
def test_probe_equals_oracle_info(name):
    cs = open(os.path.join(GOLDEN, name), "rb").read()
    i, j = oracle.info(cs), tca.probe(cs)
    assert (i.x1 - i.x0, i.y1 - i.y0, i.ncomp, i.nlevels, i.reversible, i.cbw, i.cbh, i.mct) == \
           (j.width, j.height, j.ncomp, j.nlevels, j.reversible, j.cbw, j.cbh, j.mct)


def test_jp2_wrapper_is_accepted():
    cs = open(os.path.join(GOLDEN, "openjpeg_ll_rgb64.j2k"), "rb").read()
    # minimal JP2: signature box, ftyp box, jp2h box holding only the image
    # header (64 x 64, 3 components of 8 bits, compression type 7), jp2c box
    sig = bytes([0, 0, 0, 12]) + b"jP  " + bytes([0x0D, 0x0A, 0x87, 0x0A])
    ftyp = (20).to_bytes(4, "big") + b"ftyp" + b"jp2 " + bytes(4) + b"jp2 "
    ihdr = (22).to_bytes(4, "big") + b"ihdr" + (64).to_bytes(4, "big") + (64).to_bytes(4, "big") + \
        bytes([0, 3, 7, 7, 1, 0])
    jp2h = (8 + len(ihdr)).to_bytes(4, "big") + b"jp2h" + ihdr
    jp2c = (8 + len(cs)).to_bytes(4, "big") + b"jp2c" + cs
    f = sig + ftyp + jp2h + jp2c
    assert [_key(c) for c in tca.list_cblks(f)] == [_key(c) for c in tca.list_cblks(cs)]


def test_damaged_codestreams_fail_cleanly():
    cs = open(os.path.join(GOLDEN, "openjpeg_ll_tiles_offset.j2k"), "rb").read()
    rng = np.random.default_rng(0)
    for cut in list(range(0, 200, 7)) + [len(cs) // 2, len(cs) - 3]:
        try:
            tca.list_cblks(cs[:cut])
        except tca.TCError as e:
            assert e.status in (2, 3, 4, 5)
    for _ in range(60):
        b = bytearray(cs)
        for _ in range(4):
            b[int(rng.integers(0, 200))] = int(rng.integers(0, 256))
        try:
            tca.list_cblks(bytes(b))
        except tca.TCError as e:
            assert e.status in (1, 2, 3, 4, 5, 8)


def test_unsupported_features_are_reported():
    cs = bytearray(open(os.path.join(GOLDEN, "openjpeg_ll_rgb64.j2k"), "rb").read())
    # bit depth 20 on component 0 (SIZ at offset 2, Ssiz list at 2+2+2+36+2)
    ssiz = 2 + 2 + 2 + 2 + 32 + 2  # SOC, SIZ, Lsiz, Rsiz, 8 x 32-bit sizes, Csiz
    cs[ssiz] = 19
    with pytest.raises(tca.TCError) as e:
        tca.probe(bytes(cs))
    assert e.value.status == 4


@pytest.mark.parametrize("kw", [dict(tile=(256, 256)), dict(tile=(200, 160), cbsty=5, lossy=True, bpp=12.0),
                                dict(tile=(256, 192), packed="ppt")])
def test_tiles_parsed_in_parallel_equal_oracle(kw):
    # more than a megabyte of packet data in many tiles: the parser reads the
    # tiles' packets on several threads (one per megabyte, up to eight) and
    # merges them in tile order; the table must equal the oracle's serial parse
    from thousandthchicken_amd import synth
    src = synth.image(1600, 1800, 3, 81)
    cs = synth.encode(src, levels=4, cb=32, **kw)
    assert len(cs) > (2 << 20)
    a = [_key(c) for c in oracle.list_cblks(cs)]
    b = [_key(c) for c in tca.list_cblks(cs)]
    assert a == b
