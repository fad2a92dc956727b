"""Seeded synthetic decode workloads (codestreams) of the five benchmark configurations.

  cfg1  512x512x3 8-bit, lossless RCT + 5/3, 5 levels, 64x64 blocks, 1 tile (seed 1)
  cfg2  2048x2048x3 8-bit, lossy ICT + 9/7, 5 levels, expounded quantization,
        64x64 blocks, truncated to about 1.0 bpp (seed 2)
  cfg3  3840x2160x3 12-bit, lossless RCT + 5/3, 6 levels, 1024x1024 tiles,
        32x32 blocks (seed 3)
  cfg4  256 images, sides log-uniform in [256, 2048] per axis, 1 or 3
        components, 8-bit, half lossless (RCT/5-3, 5 levels) and half lossy
        (ICT/9-7, 0.5-2 bpp), 32x32 or 64x64 blocks (image seeds 1000.., layout seed 4)
  cfg5  the cfg4 mix plus 7680x4320 16-bit single-component lossless images
        (5/3, 6 levels, 64x64 blocks), 10% of requests (seed 5)
Every codestream is made by the package's encoder; images follow synth.image.
"""
import concurrent.futures as cf
import hashlib
import math
import os

import numpy as np

from . import synth

CACHE = os.environ.get("TC_WORKLOAD_CACHE", os.path.join("/tmp", "tc_workloads"))


def _cached(key, make):
    os.makedirs(CACHE, exist_ok=True)
    path = os.path.join(CACHE, hashlib.sha1(key.encode()).hexdigest() + ".j2k")
    if os.path.exists(path):
        with open(path, "rb") as f:
            return f.read()
    data = make()
    tmp = path + f".{os.getpid()}.tmp"
    with open(tmp, "wb") as f:
        f.write(data)
    os.replace(tmp, path)
    return data


def _enc(h, w, c, bits, seed, lossless, levels, cb, bpp=None, tile=None):
    key = f"v1 {h} {w} {c} {bits} {seed} {lossless} {levels} {cb} {bpp} {tile}"

    def make():
        img = synth.image(h, w, c, seed, bits)
        return synth.encode(img, bits=bits, levels=levels, cb=cb, tile=tile, lossy=not lossless, bpp=bpp)
    return _cached(key, make)


def cfg1():
    return _enc(512, 512, 3, 8, 1, True, 5, 64)


def cfg2():
    return _enc(2048, 2048, 3, 8, 2, False, 5, 64, bpp=1.0)


def cfg3():
    return _enc(2160, 3840, 3, 12, 3, True, 6, 32, tile=(1024, 1024))


def cfg5_large(seed=5):
    return _enc(4320, 7680, 1, 16, seed, True, 6, 64)


def cfg4_layout(n=256, seed=4, smin=256, smax=2048):
    """Per-image parameters of the mixed batch (deterministic)."""
    g = synth.PCG32(synth._splitmix64(np.uint64(seed)), np.uint64(seed))
    out = []
    for i in range(n):
        u = g.uniform()[0], g.uniform()[0], g.uniform()[0], g.uniform()[0], g.uniform()[0], g.uniform()[0]
        w = int(round(math.exp(math.log(smin) + u[0] * (math.log(smax) - math.log(smin)))))
        h = int(round(math.exp(math.log(smin) + u[1] * (math.log(smax) - math.log(smin)))))
        c = 3 if u[2] < 0.5 else 1
        lossless = u[3] < 0.5
        bpp = 0.5 + 1.5 * u[4]
        cb = 32 if u[5] < 0.5 else 64
        out.append(dict(h=h, w=w, c=c, lossless=lossless, bpp=None if lossless else bpp, cb=cb, seed=1000 + i))
    return out


def cfg4(n=256, scale=1.0, threads=8):
    """The mixed batch; scale < 1 shrinks every side (for tests)."""
    lay = cfg4_layout(n)

    def one(p):
        h = max(16, int(p["h"] * scale))
        w = max(16, int(p["w"] * scale))
        return _enc(h, w, p["c"], 8, p["seed"], p["lossless"], 5, p["cb"], bpp=p["bpp"])
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, lay))


def source_image(cfg, i=0, scale=1.0):
    """The source image of a lossless workload (for bit-exact checks)."""
    if cfg == "cfg1":
        return synth.image(512, 512, 3, 1)
    if cfg == "cfg3":
        return synth.image(2160, 3840, 3, 3, 12)
    if cfg == "cfg5_large":
        return synth.image(4320, 7680, 1, 5, 16)
    if cfg == "cfg4":
        p = cfg4_layout()[i]
        return synth.image(max(16, int(p["h"] * scale)), max(16, int(p["w"] * scale)), p["c"], p["seed"])
    raise ValueError(cfg)
