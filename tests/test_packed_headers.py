"""Packed packet headers (T.800 A.7.4 PPM, A.7.5 PPT) on the CPU.

The package encoder moves each tile's packet headers, unchanged, out of the
packets into PPT segments of the tile-part header or PPM segments of the main
header (optionally split into small segments, so the reader must join them by
their index bytes, and PPM's Nppm/Ippm pairs straddle segments).  A stream and
its packed twin carry the same packets, so the oracle and the parser must give
the same image and the same code-block table for both; the in-line form is
pinned by the OpenJPEG fixtures.  No encoder available here writes packed
headers, so the marker syntax itself is pinned only by this round trip
(partially pinned)."""
import struct

import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle
from thousandthchicken_amd import synth

CASES = [
    ("rgb_l4", (150, 170, 3), dict(levels=4, cb=32)),
    ("rgb_tiles", (150, 170, 3), dict(levels=3, cb=32, tile=(64, 48))),
    ("rgb_lossy", (150, 170, 3), dict(levels=4, cb=32, lossy=True, bpp=1.0)),
    ("gray_l0", (90, 70, 1), dict(levels=0, cb=16)),
    ("gray12", (120, 100, 1), dict(levels=5, cb=64, bits=12)),
]


def _key(c):
    return (c.tile, c.comp, c.res, c.band, c.x0, c.y0, c.x1, c.y1, c.npasses, c.zero_bp, c.mb, c.nbytes, c.crc)


def _image(shape, bits=8):
    a = synth.image(shape[0], shape[1], shape[2], 17, bits)
    return a[..., 0] if shape[2] == 1 else a


def _markers(cs):
    """Marker codes of the main header and of every tile-part header."""
    out, i = [], 2
    while i + 4 <= len(cs):
        m = struct.unpack(">H", cs[i:i + 2])[0]
        if m == 0xFF90:
            psot = struct.unpack(">I", cs[i + 6:i + 10])[0]
            j = i + 12
            while struct.unpack(">H", cs[j:j + 2])[0] != 0xFF93:
                out.append(("tile", struct.unpack(">H", cs[j:j + 2])[0]))
                j += 2 + struct.unpack(">H", cs[j + 2:j + 4])[0]
            i += psot
            continue
        if m == 0xFFD9:
            break
        out.append(("main", m))
        i += 2 + struct.unpack(">H", cs[i + 2:i + 4])[0]
    return out


@pytest.mark.parametrize("packed_max", [0, 7, 100])
@pytest.mark.parametrize("packed", ["ppt", "ppm"])
@pytest.mark.parametrize("name,shape,kw", CASES, ids=[c[0] for c in CASES])
def test_packed_equals_inline(name, shape, kw, packed, packed_max):
    kw = dict(kw)
    bits = kw.pop("bits", 8)
    src = _image(shape, bits)
    inline = synth.encode(src, bits=bits, **kw)
    try:
        cs = synth.encode(src, bits=bits, packed=packed, packed_max=packed_max, **kw)
    except RuntimeError:
        pytest.skip("more than 256 segments at this payload limit")
    marks = _markers(cs)
    code = 0xFF61 if packed == "ppt" else 0xFF60
    where = "tile" if packed == "ppt" else "main"
    assert (where, code) in marks and ("main" if packed == "ppt" else "tile", code) not in marks
    if packed_max == 7:
        assert sum(1 for m in marks if m == (where, code)) > 1
    assert np.array_equal(oracle.decode(cs), oracle.decode(inline))
    a = [_key(c) for c in tca.list_cblks(cs)]
    assert a == [_key(c) for c in tca.list_cblks(inline)]
    assert a == [_key(c) for c in oracle.list_cblks(cs)]
    if not kw.get("lossy"):
        assert np.array_equal(oracle.decode(cs).transpose(1, 2, 0).reshape(src.shape), src)


def _with_main_segment(cs, seg):
    sot = cs.index(b"\xff\x90")
    return cs[:sot] + seg + cs[sot:]


def test_ppm_and_ppt_together_fail():
    src = _image((60, 70, 3))
    cs = synth.encode(src, levels=2, cb=32, packed="ppt")
    bad = _with_main_segment(cs, b"\xff\x60\x00\x07\x00\x00\x00\x00\x00")
    with pytest.raises(tca.TCError):
        tca.list_cblks(bad)
    with pytest.raises(oracle.OracleError):
        oracle.decode(bad)


def test_ppt_in_main_header_fails():
    src = _image((60, 70, 3))
    cs = synth.encode(src, levels=2, cb=32)
    bad = _with_main_segment(cs, b"\xff\x61\x00\x04\x00\x00")
    with pytest.raises(tca.TCError):
        tca.list_cblks(bad)
    with pytest.raises(oracle.OracleError):
        oracle.decode(bad)


def test_ppm_shorter_than_the_tile_parts_fails():
    src = _image((100, 100, 1))
    cs = synth.encode(src, levels=2, cb=32, tile=(50, 50), packed="ppm")
    # drop the last tile-part's Nppm/Ippm pair by shortening the one PPM segment
    i = cs.index(b"\xff\x60")
    L = struct.unpack(">H", cs[i + 2:i + 4])[0]
    pay = cs[i + 5:i + 2 + L]
    pos, pairs = 0, []
    while pos < len(pay):
        n = struct.unpack(">I", pay[pos:pos + 4])[0]
        pairs.append((pos, n))
        pos += 4 + n
    assert len(pairs) == 4
    cut = pay[:pairs[-1][0]]
    bad = cs[:i] + b"\xff\x60" + struct.pack(">H", 3 + len(cut)) + b"\x00" + cut + cs[i + 2 + L:]
    with pytest.raises(tca.TCError):
        tca.list_cblks(bad)
    with pytest.raises(oracle.OracleError):
        oracle.decode(bad)
