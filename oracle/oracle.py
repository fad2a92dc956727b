"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke test and the CPU
baseline leg of bench.py, as the checker.  The product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class Info(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("x0", "y0", "x1", "y1", "tx0", "ty0", "tw", "th", "ntx", "nty", "ncomp")] + [
        ("prec", ctypes.c_int32 * 16), ("sgnd", ctypes.c_int32 * 16)] + [
        (n, ctypes.c_int32) for n in ("nlevels", "reversible", "mct", "cbw", "cbh", "nlayers",
                                      "progression", "qstyle", "guard_bits")] + [
        ("dx", ctypes.c_int32 * 16), ("dy", ctypes.c_int32 * 16)]


class Cblk(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("tile", "comp", "res", "band", "x0", "y0", "x1", "y1", "npasses", "zero_bp", "mb",
                 "nbytes")] + [("crc", ctypes.c_uint32)]


def build():
    """Compiles liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        L.or_read_info.argtypes = [ctypes.c_char_p, ctypes.c_size_t, P(Info)]
        L.or_list_cblks.argtypes = [ctypes.c_char_p, ctypes.c_size_t, P(Cblk), ctypes.c_int]
        L.or_t1_symbol_counts.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int]
        L.or_decode.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
        L.or_decode_reduced.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_void_p]
        L.or_t1_decode.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.or_t1_decode_style.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.or_t1_decode_seg.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int] + \
            [ctypes.c_int] * 6 + [ctypes.c_void_p] * 3
        for f in ("or_idwt53", "or_fdwt53", "or_idwt97", "or_fdwt97"):
            getattr(L, f).argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 5
        L.or_mq_decode_symbols.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                                           ctypes.c_int, ctypes.c_void_p]
        L.or_tables.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.or_sign_ctx.argtypes = [ctypes.c_int, ctypes.c_int, P(ctypes.c_int)]
        _LIB = L
    return _LIB


class OracleError(RuntimeError):
    pass


def info(cs: bytes) -> Info:
    i = Info()
    rc = lib().or_read_info(cs, len(cs), ctypes.byref(i))
    if rc:
        raise OracleError(f"or_read_info failed ({rc})")
    return i


def decode(cs: bytes, want_float=False, reduce=0):
    """Full decode -> int32 array [ncomp, H, W] (and float64 pre-rounding values).
    reduce > 0 discards that many finest DWT levels: the image at the
    resolution ceil(X / 2^reduce).  With sub-sampled components the result is a
    list of per-component arrays, component c of size
    ceil(ceil(H / YRsiz) / 2^reduce) x ceil(ceil(W / XRsiz) / 2^reduce)."""
    i = info(cs)

    def c(v):
        return -((-v) >> reduce)

    def cd(v, d):
        return c(-((-v) // d))
    shape = (i.ncomp, c(i.y1) - c(i.y0), c(i.x1) - c(i.x0))
    out = np.zeros(shape, np.int32)
    outf = np.zeros(shape, np.float64) if want_float else None
    rc = lib().or_decode_reduced(cs, len(cs), reduce, out.ctypes.data,
                                 outf.ctypes.data if want_float else None)
    if rc:
        raise OracleError(f"or_decode_reduced failed ({rc})")
    if any(i.dx[k] != 1 or i.dy[k] != 1 for k in range(i.ncomp)):
        def split(a):
            res = []
            for k in range(i.ncomp):
                h = cd(i.y1, i.dy[k]) - cd(i.y0, i.dy[k])
                w = cd(i.x1, i.dx[k]) - cd(i.x0, i.dx[k])
                res.append(a[k].ravel()[:h * w].reshape(h, w).copy())
            return res
        return (split(out), split(outf)) if want_float else split(out)
    return (out, outf) if want_float else out


def list_cblks(cs: bytes):
    n = lib().or_list_cblks(cs, len(cs), None, 0)
    if n < 0:
        raise OracleError(f"or_list_cblks failed ({n})")
    arr = (Cblk * max(n, 1))()
    n2 = lib().or_list_cblks(cs, len(cs), arr, n)
    assert n2 == n
    return [arr[k] for k in range(n)]


def symbol_counts(cs: bytes):
    n = lib().or_list_cblks(cs, len(cs), None, 0)
    out = np.zeros(max(n, 1), np.int32)
    k = lib().or_t1_symbol_counts(cs, len(cs), out.ctypes.data, n)
    if k < 0:
        raise OracleError(f"or_t1_symbol_counts failed ({k})")
    return out[:n]


# Code-block style bits (T.800 Table A.19) the oracle decodes.
STYLE_BYPASS, STYLE_RESET, STYLE_TERMALL, STYLE_CAUSAL, STYLE_PTERM, STYLE_SEGSYM = 1, 2, 4, 8, 16, 32


def t1_decode(data: bytes, w, h, orient, numbps, npasses, style=0, seglen=None):
    """seglen: byte lengths of the codeword segments (BYPASS / TERMALL), in order."""
    mag = np.zeros((h, w), np.int32)
    sign = np.zeros((h, w), np.int8)
    low = np.zeros((h, w), np.int8)
    segs = np.asarray(seglen if seglen is not None else [len(data)], np.int32)
    rc = lib().or_t1_decode_seg(data, len(data), segs.ctypes.data, len(segs), w, h, orient, numbps, npasses, style,
                                mag.ctypes.data, sign.ctypes.data, low.ctypes.data)
    if rc < 0:
        raise OracleError(f"or_t1_decode failed ({rc})")
    return mag, sign, low, rc


def _dwt(name, a, x0, y0, levels, dtype):
    a = np.ascontiguousarray(a, dtype=dtype).copy()
    h, w = a.shape
    rc = getattr(lib(), name)(a.ctypes.data, w, h, x0, y0, levels)
    if rc:
        raise OracleError(f"{name} failed ({rc})")
    return a


def idwt53(a, x0=0, y0=0, levels=1):
    return _dwt("or_idwt53", a, x0, y0, levels, np.int32)


def fdwt53(a, x0=0, y0=0, levels=1):
    return _dwt("or_fdwt53", a, x0, y0, levels, np.int32)


def idwt97(a, x0=0, y0=0, levels=1):
    return _dwt("or_idwt97", a, x0, y0, levels, np.float64)


def fdwt97(a, x0=0, y0=0, levels=1):
    return _dwt("or_fdwt97", a, x0, y0, levels, np.float64)


def mq_decode_symbols(data: bytes, ctx):
    ctx = bytes(bytearray(ctx))
    out = np.zeros(len(ctx), np.uint8)
    rc = lib().or_mq_decode_symbols(data, len(data), ctx, len(ctx), out.ctypes.data)
    if rc < 0:
        raise OracleError("or_mq_decode_symbols failed")
    return out


def tables():
    zc = np.zeros((3, 256), np.uint8)
    mq = np.zeros((4, 47), np.uint32)
    lib().or_tables(zc.ctypes.data, mq.ctypes.data)
    return zc, mq


def sign_ctx(h, v):
    x = ctypes.c_int(0)
    lab = lib().or_sign_ctx(h, v, ctypes.byref(x))
    return lab, x.value
