"""The oracle against OpenJPEG 2.5.4 on codestreams that use what OpenJPEG's
encoder cannot be asked for through Pillow: the six code-block styles of
T.800 Table A.19 alone and combined (BYPASS raw segments included), PPT and
PPM packed packet headers, and sub-sampled components.  The package's encoder
wrote them; OpenJPEG's decoder, an implementation that shares no code with
this package, decoded them when the fixtures were made.  Lossless decodes must
equal OpenJPEG's exactly (and so the encoder's source), lossy ones within 1.

The raw (BYPASS) segments pin the T.800 D.6 rule that a raw sign bit is taken
as is, without the sign-context XOR of the arithmetic-coded passes: decoding a
BYPASS stream with that XOR applied differs from OpenJPEG by tens of levels."""
import numpy as np
import pytest

from oracle import oracle
from tests import pkg_fixtures as pf


@pytest.mark.parametrize("name", pf.names())
def test_oracle_equals_openjpeg_on_package_streams(name):
    cs, ref, meta = pf.load(name)
    got = oracle.decode(cs)
    d = pf.max_diff(got, ref, meta)
    assert d <= pf.tolerance(meta), (name, d)


def test_fixture_set_covers_every_style_and_both_packed_forms():
    names = pf.names()
    for sty in (1, 2, 4, 8, 16, 32):
        assert f"sty{sty}_ll" in names and f"sty{sty}_ly" in names
    assert any(n.startswith("ppt_") for n in names) and any(n.startswith("ppm_") for n in names)
    assert sum(n.startswith("sub_") for n in names) >= 3


def test_bypass_streams_hold_raw_passes():
    # raw passes start at the eleventh pass of a block (T.800 D.6, Table D.9)
    for name in ("sty1_ll", "sty1_ly", "sty1_ll_gray16", "sty63_ly"):
        cs, _, _ = pf.load(name)
        assert sum(c.npasses > 10 for c in oracle.list_cblks(cs)) >= 10, name
