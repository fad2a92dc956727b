"""Seeded synthetic images for the benchmark workloads and tests.

Recipe (per component; shape parameters drawn from a SplitMix64-seeded PCG32,
per-pixel noise from SplitMix64 applied to the pixel index):
  * 8 sinusoids a*sin(2*pi*(fx*x + fy*y) + phi), a ~ U[10,40],
    fx ~ U[0.5,16]/width, fy ~ U[0.5,16]/height, phi ~ U[0, 2*pi);
  * Gaussian noise with standard deviation 12 (the "sigma 12"), blurred with a
    Gaussian of sigma 1 pixel (7-tap separable kernel, edges clamped);
  * 20 axis-aligned sharp-edged rectangles, each side ~ U[4%, 40%] of the
    image side, that pull the area halfway towards a flat value ~ U[0,255].
Samples sum around 128, are rounded and clipped to [0,255], then scaled by
(2^B - 1)/255 and rounded for B-bit images.
"""
import math

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    with np.errstate(over="ignore"):
        x = np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)
        z = x
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


class PCG32:
    """Vector of independent PCG32 (XSH-RR) streams."""

    MUL = np.uint64(6364136223846793005)

    def __init__(self, seeds, streams):
        seeds = np.atleast_1d(np.asarray(seeds, dtype=np.uint64))
        streams = np.atleast_1d(np.asarray(streams, dtype=np.uint64))
        with np.errstate(over="ignore"):
            self.inc = (streams << np.uint64(1)) | np.uint64(1)
            self.state = np.zeros_like(seeds)
            self.next_u32()
            self.state = self.state + seeds
            self.next_u32()

    def next_u32(self):
        with np.errstate(over="ignore"):
            old = self.state
            self.state = old * self.MUL + self.inc
            xs = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)) & np.uint64(0xFFFFFFFF)
            rot = old >> np.uint64(59)
            r = ((xs >> rot) | (xs << ((np.uint64(32) - rot) & np.uint64(31)))) & np.uint64(0xFFFFFFFF)
        return r

    def uniform(self):
        """U[0,1) with 32-bit resolution, one per stream."""
        return self.next_u32().astype(np.float64) / 4294967296.0


def _scalar_stream(seed, tag):
    s = _splitmix64(np.uint64(seed) ^ _splitmix64(np.uint64(tag)))
    return PCG32(s, np.uint64(tag))


def _gauss_kernel(sigma=1.0, radius=3):
    x = np.arange(-radius, radius + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    return k / k.sum()


def _blur(a, k):
    r = len(k) // 2
    p = np.pad(a, ((0, 0), (r, r)), mode="edge")
    out = k[0] * p[:, 0:a.shape[1]]
    for i in range(1, len(k)):
        out += k[i] * p[:, i:i + a.shape[1]]
    p = np.pad(out, ((r, r), (0, 0)), mode="edge")
    out = k[0] * p[0:a.shape[0], :]
    for i in range(1, len(k)):
        out += k[i] * p[i:i + a.shape[0], :]
    return out


def component(height, width, seed, comp):
    """One float64 component of the recipe, before rounding and clipping."""
    g = _scalar_stream(seed, 1000 + comp)
    y = np.arange(height, dtype=np.float64)[:, None]
    x = np.arange(width, dtype=np.float64)[None, :]
    img = np.full((height, width), 128.0)
    for _ in range(8):
        a = 10 + 30 * g.uniform()[0]
        fx = (0.5 + 15.5 * g.uniform()[0]) / width
        fy = (0.5 + 15.5 * g.uniform()[0]) / height
        ph = 2 * math.pi * g.uniform()[0]
        # sin(u + v) = sin u cos v + cos u sin v, with u along x and v along y
        u = 2 * math.pi * fx * x + ph
        v = 2 * math.pi * fy * y
        img += a * (np.sin(u) * np.cos(v) + np.cos(u) * np.sin(v))
    for _ in range(20):
        rw = max(1, int(round(width * (0.04 + 0.36 * g.uniform()[0]))))
        rh = max(1, int(round(height * (0.04 + 0.36 * g.uniform()[0]))))
        rx = int(g.uniform()[0] * max(1, width - rw + 1))
        ry = int(g.uniform()[0] * max(1, height - rh + 1))
        val = 255 * g.uniform()[0]
        img[ry:ry + rh, rx:rx + rw] = 0.5 * img[ry:ry + rh, rx:rx + rw] + 0.5 * val
    # per-pixel noise from SplitMix64 over the pixel index (counter based), Box-Muller
    idx = np.arange(height * width, dtype=np.uint64)
    base = _splitmix64(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(comp + 1))
    with np.errstate(over="ignore"):
        z = _splitmix64(idx + base)
    u1 = ((z >> np.uint64(32)).astype(np.float64) + 1.0) / 4294967297.0
    u2 = (z & np.uint64(0xFFFFFFFF)).astype(np.float64) / 4294967296.0
    noise = (np.sqrt(-2.0 * np.log(u1)) * np.cos(2 * math.pi * u2)).reshape(height, width)
    img += _blur(12.0 * noise, _gauss_kernel(1.0))
    return img


def image(height, width, ncomp, seed, bits=8):
    """Seeded image [height, width, ncomp] (uint8 for 8-bit, else uint16)."""
    planes = [np.clip(np.rint(component(height, width, seed, c)), 0, 255) for c in range(ncomp)]
    a = np.stack(planes, axis=-1)
    if bits == 8:
        return a.astype(np.uint8)
    return np.rint(a * ((1 << bits) - 1) / 255.0).astype(np.uint16)


# ---------------------------------------------------------------- encoder

_ENC = None


def _enc_lib():
    global _ENC
    if _ENC is None:
        import ctypes
        import os
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtc_encode.so")
        if not os.path.exists(path):
            from . import build as _b
            _b.build_encoder()
        L = ctypes.CDLL(path)
        vp, i32 = ctypes.c_void_p, ctypes.c_int
        L.tce_encode.argtypes = [vp] + [i32] * 10 + [i32, ctypes.c_double, ctypes.c_double, i32, i32, i32, i32, vp,
                                                      ctypes.POINTER(ctypes.c_void_p),
                                                      ctypes.POINTER(ctypes.c_size_t)]
        L.tce_free.argtypes = [vp]
        L.tce_t1_encode_block.argtypes = [vp, i32, i32, i32, vp, i32, ctypes.POINTER(i32)]
        L.tce_t1_encode_truncated.argtypes = [vp, i32, i32, i32, i32, vp, i32, ctypes.POINTER(i32)]
        L.tce_t1_encode_truncated_style.argtypes = [vp, i32, i32, i32, i32, i32, vp, i32, ctypes.POINTER(i32)]
        L.tce_t1_encode_segments.argtypes = [vp, i32, i32, i32, i32, i32, vp, i32, ctypes.POINTER(i32), vp, i32,
                                             ctypes.POINTER(i32), ctypes.POINTER(i32)]
        _ENC = L
    return _ENC


def encode(img, bits=None, levels=5, cb=64, tile=None, mct=None, lossy=False, bpp=None, base_step=0.5,
           threads=0, cbsty=0, packed=None, packed_max=0, subsampling=None):
    """Encodes img [H, W] or [H, W, C] (uint8 / uint16) into a raw codestream.

    Lossless: RCT (when mct) + 5/3, every pass kept.  Lossy: ICT + 9/7 with
    expounded quantization, truncated to about `bpp` bits per pixel.  cbsty:
    code-block style bits written to COD and used by tier-1 (BYPASS 1, RESET 2,
    TERMALL 4, CAUSAL 8, PTERM 16, SEGSYM 32).  packed: None, "ppt" or "ppm" -- packet headers in
    the packets, in PPT segments of the tile-part headers or in PPM segments of
    the main header; packed_max caps each segment's payload (0: 65532).
    cb: code-block side, or (width, height) for non-square blocks.
    subsampling: per component (XRsiz, YRsiz); component c is then coded from
    the samples img[::YRsiz, ::XRsiz, c] (no colour transform)."""
    import ctypes
    a = np.ascontiguousarray(img)
    if a.ndim == 2:
        a = a[..., None]
    h, w, c = a.shape
    if bits is None:
        bits = 8 if a.dtype == np.uint8 else 16
    if bits > 8:
        a = np.ascontiguousarray(a.astype(np.uint16))
    else:
        a = np.ascontiguousarray(a.astype(np.uint8))
    if mct is None:
        mct = c >= 3 and not (subsampling and any(d != (1, 1) for d in subsampling))
    sub = None
    if subsampling:
        sub = np.asarray([v for d in subsampling for v in d], np.int32)
    tw, th = (tile if tile else (0, 0))
    target = (bpp * w * h / 8.0) if (lossy and bpp) else 0.0
    out = ctypes.c_void_p()
    n = ctypes.c_size_t()
    cbw, cbh = (cb, cb) if np.isscalar(cb) else cb
    cbew, cbeh = int(cbw).bit_length() - 1, int(cbh).bit_length() - 1
    rc = _enc_lib().tce_encode(a.ctypes.data, w, h, c, bits, levels, cbew, cbeh, tw, th, 1 if mct else 0,
                               1 if lossy else 0, target, base_step, threads, cbsty,
                               {None: 0, "ppt": 1, "ppm": 2}[packed], packed_max,
                               None if sub is None else sub.ctypes.data, ctypes.byref(out),
                               ctypes.byref(n))
    if rc:
        raise RuntimeError(f"tce_encode failed ({rc})")
    data = ctypes.string_at(out.value, n.value)
    _enc_lib().tce_free(out)
    return data
