"""Sub-sampled components (SIZ XRsiz/YRsiz > 1, T.800 B.2-B.3) on the CPU.

The package encoder codes component c from the samples img[::YRsiz, ::XRsiz, c]
on tile-component rectangles ceil(tile / XRsiz) (B-12).  The oracle must decode
lossless streams back to exactly those samples, component by component at its
own size, and the host parser must list the same code-blocks as the oracle
(geometry, passes, lengths, codeword hashes).  No third-party encoder here
writes sub-sampled streams, and the package encoder writes one precinct in LRCP
order, so the position-driven progression orders with sub-sampling (precinct
starts at XRsiz * 2^(PPx + levels) on the reference grid) are unpinned."""
import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle
from thousandthchicken_amd import synth

CASES = [
    ("420", (97, 131, 3), [(1, 1), (2, 2), (2, 2)], None, 8),
    ("422_tiles", (120, 100, 3), [(1, 1), (2, 1), (2, 1)], (64, 48), 8),
    ("odd_3x2", (61, 77, 1), [(3, 2)], None, 12),
    ("mixed_tiles", (100, 140, 3), [(1, 1), (1, 2), (4, 4)], (50, 50), 8),
]


def _key(c):
    return (c.tile, c.comp, c.res, c.band, c.x0, c.y0, c.x1, c.y1, c.npasses, c.zero_bp, c.mb, c.nbytes, c.crc)


@pytest.mark.parametrize("name,shape,sub,tile,bits", CASES, ids=[c[0] for c in CASES])
def test_oracle_decodes_subsampled_lossless_exactly(name, shape, sub, tile, bits):
    src = synth.image(shape[0], shape[1], shape[2], 81, bits)
    cs = synth.encode(src, bits=bits, levels=3, cb=32, subsampling=sub, tile=tile)
    got = oracle.decode(cs)
    assert isinstance(got, list) and len(got) == shape[2]
    for c, (dx, dy) in enumerate(sub):
        assert np.array_equal(got[c], src[::dy, ::dx, c]), c
    assert [_key(c) for c in tca.list_cblks(cs)] == [_key(c) for c in oracle.list_cblks(cs)]
    info = tca.probe(cs)
    assert info.planar == 1
    for c, (dx, dy) in enumerate(sub):
        assert (info.comp_height[c], info.comp_width[c]) == src[::dy, ::dx, c].shape


def test_oracle_subsampled_lossy_is_close():
    src = synth.image(97, 131, 3, 82)
    cs = synth.encode(src, levels=3, cb=32, subsampling=[(1, 1), (2, 2), (2, 2)], lossy=True, bpp=3.0)
    got = oracle.decode(cs)
    for c, d in enumerate([1, 2, 2]):
        err = np.abs(got[c].astype(np.float64) - src[::d, ::d, c])
        assert err.mean() < 6


def test_colour_transform_with_subsampling_is_rejected_by_the_encoder():
    src = synth.image(40, 40, 3, 83)
    with pytest.raises(RuntimeError):
        synth.encode(src, levels=2, cb=32, subsampling=[(1, 1), (2, 2), (2, 2)], mct=True)


def test_unsubsampled_streams_stay_interleaved():
    src = synth.image(40, 50, 3, 84)
    cs = synth.encode(src, levels=2, cb=32)
    assert tca.probe(cs).planar == 0
    assert oracle.decode(cs).shape == (3, 40, 50)
