"""The pkg_* fixtures of tests/golden/: codestreams written by this package's
encoder with code-block styles, PPT/PPM packet headers and sub-sampled
components, each with OpenJPEG 2.5.4's decode of it (tests/golden/make_fixtures.py).

OpenJPEG returns sub-sampled components upsampled by repetition, so a decode
that keeps each component at its own size is compared after the same
upsampling."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names():
    return sorted(f[4:-4] for f in os.listdir(GOLDEN) if f.startswith("pkg_") and f.endswith(".j2k"))


def load(name):
    """(codestream, OpenJPEG's decode, encoder settings)."""
    with open(os.path.join(GOLDEN, f"pkg_{name}.j2k"), "rb") as f:
        cs = f.read()
    z = np.load(os.path.join(GOLDEN, f"pkg_{name}.npz"))
    return cs, z["openjpeg"], json.loads(str(z["meta"]))


def max_diff(got, ref, meta):
    """Max |got - ref| where got is a decode as [C, H, W], [H, W, C], [H, W] or
    a list of per-component planes, and ref is OpenJPEG's [H, W(, C)]."""
    h, w, c = meta["shape"]
    ref = ref.astype(np.int64).reshape(h, w, c)
    sub = meta.get("subsampling") or [(1, 1)] * c
    if isinstance(got, list):
        planes = got
    else:
        g = np.asarray(got)
        if g.ndim == 2:
            planes = [g]
        elif g.shape[0] == c and g.shape[-1] != c:
            planes = [g[k] for k in range(c)]
        else:
            planes = [g.reshape(h, w, c)[..., k] for k in range(c)]
    d = 0
    for k, (dx, dy) in enumerate(sub):
        up = np.asarray(planes[k]).astype(np.int64)[np.arange(h)[:, None] // dy, np.arange(w)[None, :] // dx]
        d = max(d, int(np.abs(up - ref[..., k]).max()))
    return d


def tolerance(meta):
    return 1 if meta.get("lossy") else 0
