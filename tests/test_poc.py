"""Progression order changes (POC, T.800 A.6.6 and B.12.2).

The fixture openjpeg_ly_cinema4k_poc.j2k comes from OpenJPEG's Digital Cinema
4K profile: CPRL with precincts, six tile-parts and a POC marker in the first
tile-part header that runs resolutions 0-4 of every component before
resolution 5.  The generic fixture tests decode it (oracle within 1 of
OpenJPEG, parser table equal to the oracle's); the tests here show that the
POC is honoured and that the same entries placed in the main header give the
same parse."""
import os
import struct

import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAME = "openjpeg_ly_cinema4k_poc.j2k"


def _key(c):
    return (c.tile, c.comp, c.res, c.band, c.x0, c.y0, c.x1, c.y1, c.npasses, c.zero_bp, c.mb, c.nbytes, c.crc)


def _split(cs):
    """(main header, first SOT offset, POC segment offset and bytes in the first tile-part header)."""
    sot = cs.index(b"\xff\x90")
    i = sot + 12
    while cs[i:i + 2] != b"\xff\x93":
        n = struct.unpack(">H", cs[i + 2:i + 4])[0]
        if cs[i:i + 2] == b"\xff\x5f":
            return sot, i, cs[i:i + 2 + n]
        i += 2 + n
    raise AssertionError("no POC in the first tile-part header")


def _without_tile_poc(cs):
    sot, at, poc = _split(cs)
    psot = struct.unpack(">I", cs[sot + 6:sot + 10])[0]
    out = bytearray(cs[:at] + cs[at + len(poc):])
    out[sot + 6:sot + 10] = struct.pack(">I", psot - len(poc))
    return bytes(out), poc


def test_fixture_has_a_tile_header_poc_that_differs_from_cod():
    cs = open(os.path.join(GOLDEN, NAME), "rb").read()
    _, _, poc = _split(cs)
    entries = [struct.unpack(">BBHBBB", poc[4 + 7 * k:4 + 7 * k + 7]) for k in range((len(poc) - 4) // 7)]
    assert len(entries) == 2 and entries[0][3] < entries[1][3]  # two resolution ranges
    assert tca.probe(cs).progression == 4


def test_poc_is_honoured():
    cs = open(os.path.join(GOLDEN, NAME), "rb").read()
    ref = np.load(os.path.join(GOLDEN, NAME.replace(".j2k", ".npz")))["openjpeg"].astype(np.int32)
    stripped, _ = _without_tile_poc(cs)
    # without the POC the packets are read in plain CPRL order: a different parse
    assert [_key(c) for c in tca.list_cblks(stripped)] != [_key(c) for c in tca.list_cblks(cs)]
    try:
        bad = oracle.decode(stripped).transpose(1, 2, 0)
        assert np.abs(bad - ref).max() > 1
    except oracle.OracleError:
        pass


def test_main_header_poc_equals_tile_header_poc():
    cs = open(os.path.join(GOLDEN, NAME), "rb").read()
    stripped, poc = _without_tile_poc(cs)
    sot = stripped.index(b"\xff\x90")
    moved = stripped[:sot] + poc + stripped[sot:]
    assert [_key(c) for c in tca.list_cblks(moved)] == [_key(c) for c in tca.list_cblks(cs)]
    assert [_key(c) for c in oracle.list_cblks(moved)] == [_key(c) for c in oracle.list_cblks(cs)]
    assert np.array_equal(oracle.decode(moved), oracle.decode(cs))


@pytest.mark.parametrize("field,value", [(5, 7), (3, 0)])
def test_bad_poc_entries(field, value):
    """Progression 7 is rejected; an empty resolution range reads nothing
    from the POC and leaves every packet to the COD order."""
    cs = open(os.path.join(GOLDEN, NAME), "rb").read()
    sot, at, poc = _split(cs)
    b = bytearray(cs)
    offs = [4 + 7 * k for k in range((len(poc) - 4) // 7)]
    # entry layout: RSpoc, CSpoc, LYEpoc (2), REpoc, CEpoc, Ppoc
    pos = {5: 6, 3: 4}[field]
    for o in offs:
        b[at + o + pos] = value
    if field == 5:
        with pytest.raises(tca.TCError):
            tca.list_cblks(bytes(b))
        with pytest.raises(oracle.OracleError):
            oracle.decode(bytes(b))
    else:
        stripped, _ = _without_tile_poc(cs)
        assert [_key(c) for c in tca.list_cblks(bytes(b))] == [_key(c) for c in tca.list_cblks(stripped)]
        assert [_key(c) for c in oracle.list_cblks(bytes(b))] == [_key(c) for c in oracle.list_cblks(stripped)]
