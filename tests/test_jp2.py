"""The JP2 file format layer (T.800 Annex I) of the host parser, on JP2 files
OpenJPEG wrote through Pillow (tests/golden/jp2_*).

The reference reads the signature, File Type, JP2 Header and Contiguous
Codestream boxes (TC/boxes.c:142-307): it rejects a brand other than 'jp2 '
and takes the size and component count from the Image Header box.  Here, in
addition, the Image Header (and Bits Per Component) box must agree with SIZ,
and palette (pclr) and component-mapping (cmap) boxes are rejected.  Each rule
is checked on a copy of a fixture with one box changed."""
import os
import struct

import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["rgb_ll", "gray_ly", "gray16_ll"]
GEOMETRY, UNSUPPORTED, BAD_MARKER = 5, 4, 3


def load(name):
    with open(os.path.join(GOLDEN, f"jp2_{name}.jp2"), "rb") as f:
        data = f.read()
    return data, np.load(os.path.join(GOLDEN, f"jp2_{name}.npz"))


def boxes(data):
    """Top-level boxes as (type, payload) pairs."""
    out, pos = [], 0
    while pos < len(data):
        n, t = struct.unpack(">I4s", data[pos:pos + 8])
        n = n or len(data) - pos
        out.append((t, data[pos + 8:pos + n]))
        pos += n
    return out


def box(t, payload):
    return struct.pack(">I4s", 8 + len(payload), t) + payload


def build(bs):
    return b"".join(box(t, p) for t, p in bs)


def edit_header(data, fn):
    """Rebuilds the file with the JP2 Header box's sub-boxes passed through fn."""
    out = []
    for t, p in boxes(data):
        if t == b"jp2h":
            p = build(fn(boxes(p)))
        out.append((t, p))
    return build(out)


def status_of(data):
    try:
        tca.probe(data)
        return 0
    except tca.TCError as e:
        return e.status


def raw_codestream(data):
    return dict(boxes(data))[b"jp2c"]


@pytest.mark.parametrize("name", NAMES)
def test_jp2_fixture_probes_like_its_codestream(name):
    data, _ = load(name)
    a, b = tca.probe(data), tca.probe(raw_codestream(data))
    assert a.jp2 == 1 and b.jp2 == 0
    assert a.colourspace == (16 if name.startswith("rgb") else 17)
    for f in ("width", "height", "ncomp", "max_prec", "reversible", "nlevels", "cbw", "cbh"):
        assert getattr(a, f) == getattr(b, f), f


@pytest.mark.parametrize("name", NAMES)
def test_oracle_decodes_jp2_fixture(name):
    data, z = load(name)
    out = oracle.decode(data)
    got = out[0] if z["source"].ndim == 2 else out.transpose(1, 2, 0)
    if name.endswith("_ll"):
        assert np.array_equal(got, z["source"]) and np.array_equal(got, z["openjpeg"])
    else:
        assert np.abs(got.astype(np.int32) - z["openjpeg"]).max() <= 1


def _ihdr(fn):
    def f(sub):
        return [(t, fn(bytearray(p)) if t == b"ihdr" else p) for t, p in sub]
    return f


def _set(off, fmt, v):
    def f(p):
        struct.pack_into(fmt, p, off, v)
        return bytes(p)
    return f


def test_image_header_must_agree_with_siz():
    data, _ = load("rgb_ll")
    h, w = struct.unpack(">II", dict(boxes(dict(boxes(data))[b"jp2h"]))[b"ihdr"][:8])
    assert (w, h) == (64, 48)
    assert status_of(edit_header(data, _ihdr(_set(4, ">I", w + 1)))) == GEOMETRY
    assert status_of(edit_header(data, _ihdr(_set(0, ">I", h - 1)))) == GEOMETRY
    assert status_of(edit_header(data, _ihdr(_set(8, ">H", 1)))) == GEOMETRY    # components
    assert status_of(edit_header(data, _ihdr(_set(10, ">B", 9)))) == GEOMETRY   # 10-bit
    assert status_of(edit_header(data, _ihdr(_set(10, ">B", 0x87)))) == GEOMETRY  # signed
    assert status_of(edit_header(data, _ihdr(_set(11, ">B", 8)))) == UNSUPPORTED  # compression type


def test_bits_per_component_box():
    data, _ = load("rgb_ll")
    def with_bpcc(vals):
        def f(sub):
            sub = [(t, _set(10, ">B", 255)(bytearray(p)) if t == b"ihdr" else p) for t, p in sub]
            return sub[:1] + [(b"bpcc", bytes(vals))] + sub[1:]
        return f
    assert status_of(edit_header(data, with_bpcc([7, 7, 7]))) == 0
    assert status_of(edit_header(data, with_bpcc([7, 7, 11]))) == GEOMETRY
    assert status_of(edit_header(data, with_bpcc([7, 7]))) == GEOMETRY


@pytest.mark.parametrize("kind", [b"pclr", b"cmap"])
def test_palette_and_component_mapping_are_rejected(kind):
    data, _ = load("rgb_ll")
    payload = {b"pclr": struct.pack(">HB", 2, 1) + bytes([7]) + bytes([0, 255]),
               b"cmap": struct.pack(">HBB", 0, 1, 0)}[kind]
    bad = edit_header(data, lambda sub: sub + [(kind, payload)])
    assert status_of(bad) == UNSUPPORTED
    assert kind.decode() in tca.lib().tc_last_error().decode()


def test_file_type_brand():
    data, _ = load("gray16_ll")
    def with_ftyp(payload):
        return build([(t, payload if t == b"ftyp" else p) for t, p in boxes(data)])
    assert status_of(with_ftyp(b"jpx \x00\x00\x00\x00jpx ")) == UNSUPPORTED
    # another brand is fine when the compatibility list names 'jp2 ' (I.5.2)
    assert status_of(with_ftyp(b"jpx \x00\x00\x00\x00jpx jp2 ")) == 0


def test_box_order_and_presence():
    data, _ = load("gray_ly")
    bs = boxes(data)
    by = dict(bs)
    sig, ftyp = bs[0], bs[1]
    # JP2 Header after the codestream, or missing
    assert status_of(build([sig, ftyp, (b"jp2c", by[b"jp2c"]), (b"jp2h", by[b"jp2h"])])) == BAD_MARKER
    assert status_of(build([sig, ftyp, (b"jp2c", by[b"jp2c"])])) == BAD_MARKER
    # File Type box not right after the signature
    assert status_of(build([sig, (b"jp2h", by[b"jp2h"]), ftyp, (b"jp2c", by[b"jp2c"])])) == BAD_MARKER
    # Image Header not the first box of the JP2 Header
    assert status_of(edit_header(data, lambda sub: sub[1:] + sub[:1])) == BAD_MARKER
    # boxes this reader does not use are skipped
    extra = build([sig, ftyp, (b"xml ", b"<a/>"), (b"jp2h", by[b"jp2h"]), (b"uuid", bytes(20)),
                   (b"jp2c", by[b"jp2c"])])
    assert status_of(extra) == 0
    assert np.array_equal(oracle.decode(extra), oracle.decode(data))


def test_codestream_box_to_end_of_file():
    # a last box of length 0 runs to the end of the file (I.4)
    data, _ = load("rgb_ll")
    bs = boxes(data)
    head = build(bs[:-1])
    cs = dict(bs)[b"jp2c"]
    assert status_of(head + struct.pack(">I4s", 0, b"jp2c") + cs) == 0
