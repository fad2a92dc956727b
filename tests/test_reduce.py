"""Resolution-reduced decode on the CPU.

The oracle's reduced decode of a lossless stream is pinned by a closed form
that shares no code with the decoder's inverse path: the level-k low-pass band
of the source, computed with the forward transforms.  The encoder applies the
DC shift, the forward RCT and an L-level forward 5/3 DWT; undoing L - k of
those levels leaves exactly the LL band of a k-level forward DWT, which the
reduced decoder then passes through the inverse RCT and the level shift
(T.800 B.5, F.2, G.2).  The reduced sizes reported by the C ABI's probe are
checked against ceil(X / 2^k) for the image and tile origins the encoder
writes."""
import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle
from thousandthchicken_amd import synth


def _cdiv(v, k):
    return -((-v) >> k)


def _rct(a):
    r, g, b = (a[..., i].astype(np.int64) for i in range(3))
    return [(r + 2 * g + b) >> 2, b - g, r - g]


def _irct(y, u, v):
    g = y - ((u + v) >> 2)
    return np.stack([v + g, g, u + g])


def _expected(src, bits, k, mct):
    """Level-k low-pass image of src (closed form described above)."""
    h, w = src.shape[:2]
    x = src.astype(np.int64) - (1 << (bits - 1))
    comps = _rct(x) if mct else [x[..., c] for c in range(x.shape[-1])]
    ll = []
    for c in comps:
        t = np.asarray(oracle.fdwt53(c.astype(np.int32).copy(), 0, 0, k))
        ll.append(t[:_cdiv(h, k), :_cdiv(w, k)].astype(np.int64))
    out = _irct(*ll) if mct else np.stack(ll)
    return np.clip(out + (1 << (bits - 1)), 0, (1 << bits) - 1)


@pytest.mark.parametrize("shape,ncomp,bits,levels,cb", [
    ((97, 131), 3, 8, 4, 32),
    ((64, 64), 1, 8, 5, 64),
    ((45, 200), 1, 12, 3, 32),
    ((130, 77), 3, 10, 5, 64),
])
def test_oracle_reduced_lossless_equals_low_pass_band(shape, ncomp, bits, levels, cb):
    src = synth.image(shape[0], shape[1], ncomp, 31 + levels, bits)
    cs = synth.encode(src, bits=bits, levels=levels, cb=cb)
    mct = ncomp == 3
    assert np.array_equal(oracle.decode(cs).transpose(1, 2, 0), src.reshape(shape + (ncomp,)))
    for k in range(1, levels + 1):
        got = oracle.decode(cs, reduce=k)
        exp = _expected(src.reshape(shape + (ncomp,)), bits, k, mct)
        assert got.shape == exp.shape, k
        assert np.array_equal(got, exp), k


def test_oracle_reduce_zero_is_full_decode():
    src = synth.image(50, 70, 3, 4)
    cs = synth.encode(src, levels=3, cb=32, lossy=True, bpp=1.5)
    assert np.array_equal(oracle.decode(cs, reduce=0), oracle.decode(cs))


def test_reduce_beyond_levels_fails():
    cs = synth.encode(synth.image(40, 40, 1, 2), levels=2, cb=32)
    with pytest.raises(oracle.OracleError):
        oracle.decode(cs, reduce=3)
    with pytest.raises(tca.TCError):
        tca.probe(cs, reduce=3)


@pytest.mark.parametrize("levels", [1, 3, 5])
def test_probe_reduced_sizes(levels):
    cs = synth.encode(synth.image(97, 131, 3, 5), levels=levels, cb=32)
    for k in range(levels + 1):
        i = tca.probe(cs, reduce=k)
        o = oracle.decode(cs, reduce=k)
        assert (i.height, i.width) == o.shape[1:] == (_cdiv(97, k), _cdiv(131, k))
        assert i.nlevels == levels - k
