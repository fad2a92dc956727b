"""Seeded test images and codestreams for the tests.

Images follow the benchmark recipe in small sizes; codestreams come from the
committed fixtures in tests/golden/ or, for extra cases, from OpenJPEG through
Pillow when it is importable.
"""
import io
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def pillow_j2k():
    try:
        from PIL import Image, features
        return Image if features.check("jpg_2000") else None
    except Exception:
        return None


def encode_openjpeg(a, **kw):
    """Encodes with OpenJPEG (via Pillow) into a raw .j2k codestream."""
    Image = pillow_j2k()
    if Image is None:
        raise RuntimeError("Pillow with OpenJPEG is not available")
    b = io.BytesIO()
    if a.ndim == 2 and a.dtype == np.uint16:
        Image.fromarray(a, "I;16").save(b, "JPEG2000", no_jp2=True, **kw)
    else:
        Image.fromarray(a, "RGB" if a.ndim == 3 else "L").save(b, "JPEG2000", no_jp2=True, **kw)
    return b.getvalue()


def decode_openjpeg(cs):
    Image = pillow_j2k()
    im = Image.open(io.BytesIO(cs))
    a = np.asarray(im)
    return a


def load_fixture(name):
    with open(os.path.join(GOLDEN, name), "rb") as f:
        return f.read()
