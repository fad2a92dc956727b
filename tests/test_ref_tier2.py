"""Tier-2 parity against the reference itself: the reference's own tier-2
parser (TC/codestream.c, compiled from its sources by `make -C oracle ref`
into oracle/_ref/tier2_ref) and this repository's parser must list the same
code-blocks -- position, pass count, significant bit-planes, codeword length
and a hash of the codeword bytes -- for the same codestream.  The reference
handles raw, reversible codestreams with one tile-part per tile, one layer and
one precinct, and no COM marker, so COM segments are removed first.  It also
sizes every tile from its encoder configuration instead of SIZ (tile = whole
image unless configured, TC/codestream_image.c:242-243), so multi-tile
streams are compared only to show that defect; their tier-2 parity rests on
the OpenJPEG fixtures (tests/test_parser.py)."""
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

import thousandthchicken_amd as tca
from thousandthchicken_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "tier2_ref")
GOLDEN = os.path.join(ROOT, "tests", "golden")

pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="reference tier-2 build absent (make -C oracle ref)")


def strip_com(cs: bytes) -> bytes:
    """Drops COM segments of the main header (the reference does not skip them)."""
    out = bytearray(cs[:2])
    i = 2
    while i + 4 <= len(cs):
        m = struct.unpack(">H", cs[i:i + 2])[0]
        if m == 0xFF90:  # first SOT: the rest is copied unchanged
            out += cs[i:]
            return bytes(out)
        n = struct.unpack(">H", cs[i + 2:i + 4])[0]
        if m != 0xFF64:
            out += cs[i:i + 2 + n]
        i += 2 + n
    return bytes(out)


def ref_table(cs: bytes):
    with tempfile.NamedTemporaryFile(suffix=".j2k") as f:
        f.write(cs)
        f.flush()
        r = subprocess.run([REF, f.name], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    rows = []
    for line in r.stdout.split("\n"):
        v = line.split()
        if len(v) != 12:
            continue
        t, c, res, o, x0, y0, x1, y1, np_, sb, ln, crc = map(int, v)
        rows.append((t, c, res, o, x1 - x0, y1 - y0, np_, sb if np_ else 0, ln, crc))
    return sorted(rows)


def our_table(cs: bytes):
    rows = []
    for b in tca.list_cblks(cs):
        rows.append((b.tile, b.comp, b.res, b.band, b.x1 - b.x0, b.y1 - b.y0, b.npasses,
                     (b.mb - b.zero_bp) if b.npasses else 0, b.nbytes, b.crc))
    return sorted(rows)


def _cases():
    cases = []
    for shape, bits, kw in [((128, 160, 3), 8, dict(levels=5, cb=64)),
                            ((200, 136, 3), 12, dict(levels=4, cb=32)),
                            ((96, 130, 1), 16, dict(levels=3, cb=32)),
                            ((67, 45, 1), 8, dict(levels=2, cb=16)),
                            ((300, 257, 3), 10, dict(levels=6, cb=64))]:
        img = synth.image(shape[0], shape[1], shape[2], 21, bits)
        cases.append((f"encoder {shape} {bits}-bit", synth.encode(img, bits=bits, **kw)))
    for name in sorted(os.listdir(GOLDEN)):
        if name.startswith("openjpeg_ll_") and name.endswith(".j2k") and "rpcl" not in name and "tiles" not in name \
                and "prec" not in name:
            cases.append((name, strip_com(open(os.path.join(GOLDEN, name), "rb").read())))
    return cases


@pytest.mark.parametrize("name,cs", _cases(), ids=lambda v: v if isinstance(v, str) else "")
def test_tier2_tables_match_reference(name, cs):
    ours, ref = our_table(cs), ref_table(cs)
    assert len(ours) == len(ref), (len(ours), len(ref))
    assert ours == ref


def test_reference_sizes_tiles_from_the_whole_image():
    """The reference's tile-size defect on a multi-tile stream: the same packet
    bytes, but tile 0's LL block takes the whole image's LL size."""
    img = synth.image(200, 136, 3, 21, 12)
    cs = synth.encode(img, bits=12, levels=4, cb=32, tile=(64, 96))
    ref = [r for r in ref_table(cs) if r[:4] == (0, 0, 0, 0)]
    ours = [r for r in our_table(cs) if r[:4] == (0, 0, 0, 0)]
    assert ours[0][4:6] == (4, 6)             # 64x96 tile, 4 levels
    assert ref[0][4:6] == (9, 13)             # 136x200 image, 4 levels
    assert ref[0][8:] == ours[0][8:]          # same length and codeword bytes
