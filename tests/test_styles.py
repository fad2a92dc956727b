"""Code-block styles (T.800 Table A.19) on the CPU: BYPASS (1), RESET (2),
TERMALL (4), CAUSAL (8), PTERM (16) and SEGSYM (32) in the oracle and the host
parser.

OpenJPEG's encoder cannot be asked for these styles through Pillow, but its
decoder reads them: tests/test_openjpeg_pins.py checks the oracle against
OpenJPEG's decode of styled codestreams written by this package's encoder
(every style alone and combined, BYPASS raw segments included).  Here the
styled tier-1 is tied, block by block, to the style-free decode:
- every full-pass styled block decodes to its source coefficients;
- BYPASS, RESET, TERMALL, PTERM and SEGSYM change how passes are coded (raw
  bits, context resets, terminated segments), never which bits a pass carries,
  so at every truncation point a styled block decodes exactly as the
  style-free encoding of the same coefficients; CAUSAL changes which
  coefficients the significance pass visits, so for it the equality holds at
  truncation points that end on a cleanup pass; BYPASS and TERMALL truncate
  only where a codeword segment ends;
- decoding a CAUSAL stream without the flag, or a style-free stream with
  SEGSYM, fails (wrong coefficients or a corrupt-segment error), so the flags
  are not ignored.
- a BYPASS or TERMALL block decoded as one segment, without its segment
  lengths, decodes wrongly."""
import ctypes

import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle
from thousandthchicken_amd import synth

STYLES = [1, 2, 4, 8, 16, 32, 1 | 4, 2 | 8 | 32, 2 | 8 | 16 | 32, 63]


def _encode(a, orient, keep, style):
    """(codeword, segment lengths, bit-planes, passes kept) of one block."""
    L = synth._enc_lib()
    h, w = a.shape
    buf = np.zeros(1 << 18, np.uint8)
    segs = np.zeros(128, np.uint32)
    nb, ns, npas = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    k = L.tce_t1_encode_segments(a.ctypes.data, w, h, orient, keep, style, buf.ctypes.data, buf.size,
                                 ctypes.byref(nb), segs.ctypes.data, len(segs), ctypes.byref(ns), ctypes.byref(npas))
    assert k >= 0
    return bytes(buf[:k]), [int(v) for v in segs[:ns.value]], nb.value, npas.value


def styled_blocks(seed, n, style, maxw=64, maxh=64, full=False):
    """Random code-blocks encoded with `style`, each with the style-free
    encoding of the same coefficients and truncation point."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        w, h = int(rng.integers(1, maxw + 1)), int(rng.integers(1, maxh + 1))
        orient = int(rng.integers(0, 4))
        bits = int(rng.integers(1, 16))
        a = rng.integers(-(1 << bits) + 1, 1 << bits, (h, w)).astype(np.int32)
        a[rng.uniform(size=(h, w)) < rng.uniform(0, 0.95)] = 0
        keep = 999 if full or rng.uniform() < 0.4 else int(rng.integers(1, 50))
        data, seglen, numbps, npass = _encode(a, orient, keep, style)
        if numbps == 0 or npass == 0:
            continue
        plain, _, nb0, np0 = _encode(a, orient, npass, 0)
        assert nb0 == numbps and np0 == npass
        out.append(dict(w=w, h=h, orient=orient, numbps=numbps, npasses=npass, data=data, seglen=seglen,
                        plain=plain, src=a, style=style))
    return out


def reconstruct(mag, sign, low):
    half = np.where(low >= 1, np.left_shift(1, np.maximum(low.astype(np.int32) - 1, 0)), 0)
    v = np.where(low >= 0, mag + half, 0)
    return np.where(sign.astype(bool), -v, v).astype(np.int32)


def _decode(b, data, style, seglen=None):
    return reconstruct(*oracle.t1_decode(data, b["w"], b["h"], b["orient"], b["numbps"], b["npasses"], style,
                                         seglen)[:3])


@pytest.mark.parametrize("style", STYLES)
def test_oracle_styled_full_blocks_round_trip(style):
    blocks = styled_blocks(100 + style, 60, style, full=True)
    assert len(blocks) > 40
    for b in blocks:
        assert np.array_equal(_decode(b, b["data"], style, b["seglen"]), b["src"]), (b["w"], b["h"], b["orient"])


@pytest.mark.parametrize("style", STYLES)
def test_oracle_styled_truncated_equals_style_free(style):
    blocks = styled_blocks(200 + style, 120, style)
    checked = 0
    for b in blocks:
        if (style & oracle.STYLE_CAUSAL) and b["npasses"] % 3 != 1:
            continue
        got = _decode(b, b["data"], style, b["seglen"])
        exp = _decode(b, b["plain"], 0)
        assert np.array_equal(got, exp), (b["w"], b["h"], b["orient"], b["numbps"], b["npasses"])
        checked += 1
    assert checked > 20


def test_causal_flag_changes_the_decode():
    blocks = styled_blocks(300, 40, oracle.STYLE_CAUSAL, maxw=32, maxh=32, full=True)
    wrong = 0
    for b in blocks:
        if b["h"] <= 4:
            continue
        try:
            wrong += not np.array_equal(_decode(b, b["data"], 0), b["src"])
        except oracle.OracleError:
            wrong += 1
    assert wrong >= 10


def test_segsym_rejects_streams_without_the_marker():
    blocks = styled_blocks(301, 40, 0, full=True)
    bad = 0
    for b in blocks:
        with pytest.raises(oracle.OracleError):
            oracle.t1_decode(b["data"], b["w"], b["h"], b["orient"], b["numbps"], b["npasses"], oracle.STYLE_SEGSYM)
        bad += 1
    assert bad > 20


@pytest.mark.parametrize("style", [1, 4, 1 | 4])
def test_segments_are_needed_for_bypass_and_termall(style):
    """Decoded as one segment (no segment lengths), a BYPASS / TERMALL block
    goes wrong: its passes restart the decoder or are raw."""
    blocks = styled_blocks(400 + style, 40, style, full=True)
    wrong = 0
    for b in [b for b in blocks if len(b["seglen"]) > 1]:
        try:
            wrong += not np.array_equal(_decode(b, b["data"], style), b["src"])
        except oracle.OracleError:
            wrong += 1
        if style & 4:
            assert len(b["seglen"]) == b["npasses"]
    assert wrong >= 10


@pytest.mark.parametrize("style", STYLES)
@pytest.mark.parametrize("lossy", [False, True])
def test_styled_codestreams_decode_in_oracle_and_parse_alike(style, lossy):
    src = synth.image(90, 110, 3, 7 + style)
    cs = synth.encode(src, levels=3, cb=32, cbsty=style, lossy=lossy, bpp=2.0)
    got = oracle.decode(cs).transpose(1, 2, 0)
    if not lossy:
        assert np.array_equal(got, src)
    else:
        # rate control keeps different passes per style (BYPASS / TERMALL only at
        # segment ends): check the fidelity only
        mse = np.mean((got.astype(np.float64) - src) ** 2)
        assert 10 * np.log10(255.0 ** 2 / mse) > (20 if style & 5 else 25)
    key = lambda c: (c.tile, c.comp, c.res, c.band, c.x0, c.y0, c.x1, c.y1, c.npasses, c.zero_bp, c.mb, c.nbytes,
                     c.crc)
    assert [key(c) for c in tca.list_cblks(cs)] == [key(c) for c in oracle.list_cblks(cs)]


def _cod_style_offset(cs):
    i = cs.index(b"\xff\x52")
    # marker, Lcod, Scod, progression, layers (2), MCT, levels, xcb, ycb, style
    return i + 2 + 2 + 1 + 1 + 2 + 1 + 1 + 1 + 1


@pytest.mark.parametrize("flag", [64, 128])
def test_reserved_style_bits_are_reported_unsupported(flag):
    cs = synth.encode(synth.image(40, 40, 1, 3), levels=2, cb=32, cbsty=2)
    off = _cod_style_offset(cs)
    assert cs[off] == 2
    b = bytearray(cs)
    b[off] |= flag
    with pytest.raises(tca.TCError) as e:
        tca.probe(bytes(b))
    assert e.value.status == 4
    with pytest.raises(oracle.OracleError):
        oracle.decode(bytes(b))
