"""Writes the committed golden fixtures of tests/golden/.

1. reference_tables.json: the constant tables the reference holds
   (TC/coefficient_coder.cl:36-60 MQ tables, :330-415 SPCXLUT, :437-451
   signcxlut; TC/dwt_r97.cl:5-8 and TC/dwt_common.cl:10-11 9/7 constants),
   parsed as numbers from the reference sources.  Run in the authoring
   container, where /root/reference exists.
2. openjpeg_*.j2k with openjpeg_*.npz: small codestreams coded by OpenJPEG
   2.5.4 (bundled with Pillow 12.2.0) from seeded images, with the source
   image and OpenJPEG's own decode of each, an independent decoder's output
   for the oracle and the GPU path to match.
3. pkg_*.j2k with pkg_*.npz: codestreams written by this package's encoder
   with the features OpenJPEG's encoder cannot be asked for through Pillow
   (code-block styles BYPASS/RESET/TERMALL/CAUSAL/PTERM/SEGSYM alone and
   combined, PPT and PPM packed packet headers, sub-sampled components),
   each with OpenJPEG's decode of it ("openjpeg") and the encoder settings
   ("meta", JSON).  OpenJPEG returns sub-sampled components upsampled to the
   full grid by repetition (sample (x, y) of component c is component c's
   sample (x // XRsiz, y // YRsiz)); the layouts chosen avoid the one
   OpenJPEG treats as sYCC (full-size first component, sub-sampled second and
   third), which it colour-converts.
4. jp2_*.jp2 with jp2_*.npz: JP2 files (signature, File Type, JP2 Header with
   Image Header and Colour Specification, Contiguous Codestream boxes) written
   by OpenJPEG through Pillow, with the source image and OpenJPEG's decode.
"""
import io
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = "/root/reference/ThousandthChicken"


def c_array(text, name):
    m = re.search(name + r"[^=]*=\s*\{(.*?)\};", text, re.S)
    body = re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S)
    return [int(v, 0) for v in re.findall(r"0x[0-9A-Fa-f]+|\d+", body)]


def tables():
    cc = open(os.path.join(REF, "coefficient_coder.cl")).read()
    out = {
        "source": "TC/coefficient_coder.cl, TC/dwt_r97.cl, TC/dwt_common.cl",
        "Qe": c_array(cc, r"unsigned short Qe\[\]"),
        "NMPS": c_array(cc, r"unsigned char NMPS\[\]"),
        "NLPS": c_array(cc, r"unsigned char NLPS\[\]"),
        "SWITCH": c_array(cc, r"unsigned char SWITCH\[\]"),
        "signcxlut": c_array(cc, r"unsigned char signcxlut\[256\]"),
    }
    sp = c_array(cc, r"unsigned char SPCXLUT\[3\]\[512\]")
    out["SPCXLUT"] = [sp[0:512], sp[512:1024], sp[1024:1536]]
    r97 = open(os.path.join(REF, "dwt_r97.cl")).read() + open(os.path.join(REF, "dwt_common.cl")).read()
    consts = {}
    for m in re.finditer(r"CONSTANT float (\w+)\s*=\s*([-+]?\d*\.\d+)f?\s*;", r97):
        consts[m.group(1)] = float(m.group(2))
    out["dwt97_constants"] = consts
    assert len(out["Qe"]) == 47 and len(out["signcxlut"]) == 256
    return out


CASES = [
    # name, (h, w, c, bits), seed, pillow options
    ("ll_rgb64", (64, 64, 3, 8), 21, dict(num_resolutions=3, mct=1)),
    ("ll_rgb100x77_cb32", (100, 77, 3, 8), 22, dict(num_resolutions=4, codeblock_size=(32, 32), mct=1)),
    ("ll_tiles_offset", (150, 110, 3, 8), 23, dict(num_resolutions=4, tile_size=(64, 64), offset=(3, 5), mct=1)),
    ("ll_rpcl", (97, 131, 3, 8), 24, dict(num_resolutions=5, progression="RPCL", mct=1)),
    ("ll_gray16", (120, 90, 1, 16), 25, dict(num_resolutions=5)),
    ("ll_gray8_nomct", (70, 50, 1, 8), 26, dict(num_resolutions=2)),
    ("ly_rgb128", (128, 128, 3, 8), 27, dict(irreversible=True, quality_mode="rates", quality_layers=[24],
                                               num_resolutions=5, mct=1)),
    ("ly_rgb_layers", (111, 143, 3, 8), 28, dict(irreversible=True, quality_mode="rates",
                                                   quality_layers=[40, 20, 8], num_resolutions=4, mct=1)),
    ("ly_gray", (96, 100, 1, 8), 29, dict(irreversible=True, quality_mode="rates", quality_layers=[6],
                                          num_resolutions=4)),
    # precincts (OpenJPEG halves the given size at each lower resolution)
    ("ll_prec_pcrl_tiles", (90, 120, 3, 8), 30, dict(num_resolutions=4, codeblock_size=(16, 16),
                                                      precinct_size=(64, 64), progression="PCRL",
                                                      tile_size=(64, 64), offset=(3, 5), mct=1)),
    ("ll_prec_cprl", (70, 90, 3, 8), 31, dict(num_resolutions=4, codeblock_size=(16, 16), precinct_size=(32, 32),
                                               progression="CPRL", mct=1)),
    ("ly_prec_rpcl_layers", (100, 100, 3, 8), 32, dict(irreversible=True, quality_mode="rates",
                                                        quality_layers=[30, 10], num_resolutions=4,
                                                        codeblock_size=(16, 16), precinct_size=(64, 64),
                                                        progression="RPCL", mct=1)),
    # Digital Cinema profiles: CPRL with precincts, tile-parts per resolution
    # range and TLM; the 4K profile writes a POC marker in the tile header
    ("ly_cinema4k_poc", (200, 256, 3, 8), 33, dict(cinema_mode="cinema4k-24")),
    ("ly_cinema2k", (200, 256, 3, 8), 34, dict(cinema_mode="cinema2k-24")),
]


def codestreams():
    from PIL import Image
    from thousandthchicken_amd import synth
    for name, (h, w, c, bits), seed, kw in CASES:
        a = synth.image(h, w, c, seed, bits)
        src = a[..., 0] if c == 1 else a
        b = io.BytesIO()
        if bits == 16:
            Image.fromarray(src, "I;16").save(b, "JPEG2000", no_jp2=True, **kw)
        else:
            Image.fromarray(src, "RGB" if c == 3 else "L").save(b, "JPEG2000", no_jp2=True, **kw)
        cs = b.getvalue()
        dec = np.asarray(Image.open(io.BytesIO(cs)))
        with open(os.path.join(HERE, f"openjpeg_{name}.j2k"), "wb") as f:
            f.write(cs)
        np.savez_compressed(os.path.join(HERE, f"openjpeg_{name}.npz"), source=src, openjpeg=dec)
        print(name, len(cs))


# name, (h, w, c, bits), seed, synth.encode options
_PKG_IMG = (64, 72, 3, 8)
PKG_CASES = []
for _sty in (1, 2, 4, 8, 16, 32, 1 | 4, 1 | 8, 63):
    PKG_CASES.append((f"sty{_sty}_ll", _PKG_IMG, 51, dict(levels=3, cb=16, cbsty=_sty)))
    PKG_CASES.append((f"sty{_sty}_ly", _PKG_IMG, 51, dict(levels=3, cb=16, cbsty=_sty, lossy=True, bpp=12.0)))
PKG_CASES += [
    # 16-bit samples: more bit-planes, so most passes of most blocks are raw under BYPASS
    ("sty1_ll_gray16", (56, 60, 1, 16), 52, dict(levels=3, cb=32, cbsty=1)),
    ("sty63_ll_gray16", (56, 60, 1, 16), 52, dict(levels=3, cb=32, cbsty=63)),
    ("ppt_ll", _PKG_IMG, 53, dict(levels=3, cb=16, packed="ppt", packed_max=100)),
    ("ppt_ly_tiles", _PKG_IMG, 53, dict(levels=3, cb=16, packed="ppt", packed_max=100, tile=(40, 32),
                                        lossy=True, bpp=2.0)),
    ("ppm_ll_tiles", _PKG_IMG, 53, dict(levels=3, cb=16, packed="ppm", packed_max=100, tile=(40, 32))),
    ("ppm_ly", _PKG_IMG, 53, dict(levels=3, cb=16, packed="ppm", packed_max=100, lossy=True, bpp=2.0)),
    ("ppm_sty63_ly_tiles", _PKG_IMG, 54, dict(levels=3, cb=16, packed="ppm", packed_max=64, tile=(40, 32),
                                              cbsty=63, lossy=True, bpp=12.0)),
    ("sub_21_11_12_ll", _PKG_IMG, 55, dict(levels=3, cb=16, subsampling=[(2, 1), (1, 1), (1, 2)])),
    ("sub_12_21_22_ll_tiles", _PKG_IMG, 55, dict(levels=3, cb=16, subsampling=[(1, 2), (2, 1), (2, 2)],
                                                 tile=(40, 32))),
    ("sub_22_ly", _PKG_IMG, 55, dict(levels=3, cb=16, subsampling=[(2, 2)] * 3, lossy=True, bpp=3.0)),
]


def package_streams():
    from PIL import Image
    from thousandthchicken_amd import synth
    for name, (h, w, c, bits), seed, kw in PKG_CASES:
        a = synth.image(h, w, c, seed, bits)
        src = a[..., 0] if c == 1 else a
        cs = synth.encode(src, bits=bits, **kw)
        dec = np.asarray(Image.open(io.BytesIO(cs)))
        if not kw.get("lossy"):
            sub = kw.get("subsampling") or [(1, 1)] * c
            for k, (dx, dy) in enumerate(sub):
                plane = dec if c == 1 else dec[..., k]
                comp = src if c == 1 else src[::dy, ::dx, k]
                up = comp[np.arange(h)[:, None] // dy, np.arange(w)[None, :] // dx]
                assert np.array_equal(plane, up), (name, k)
        meta = dict(kw, shape=[h, w, c], bits=bits, seed=seed)
        with open(os.path.join(HERE, f"pkg_{name}.j2k"), "wb") as f:
            f.write(cs)
        np.savez_compressed(os.path.join(HERE, f"pkg_{name}.npz"), openjpeg=dec, meta=np.array(json.dumps(meta)))
        print(name, len(cs))


# JP2 files (T.800 Annex I boxes) written by OpenJPEG through Pillow: name,
# (h, w, c, bits), seed, Pillow options
JP2_CASES = [
    ("rgb_ll", (48, 64, 3, 8), 56, dict(num_resolutions=3, mct=1)),
    ("gray_ly", (40, 56, 1, 8), 57, dict(num_resolutions=3, irreversible=True, quality_mode="rates",
                                         quality_layers=[8])),
    ("gray16_ll", (36, 44, 1, 16), 58, dict(num_resolutions=3)),
]


def jp2_files():
    from PIL import Image
    from thousandthchicken_amd import synth
    for name, (h, w, c, bits), seed, kw in JP2_CASES:
        a = synth.image(h, w, c, seed, bits)
        src = a[..., 0] if c == 1 else a
        b = io.BytesIO()
        mode = "I;16" if bits == 16 else ("RGB" if c == 3 else "L")
        Image.fromarray(src, mode).save(b, "JPEG2000", no_jp2=False, **kw)
        data = b.getvalue()
        dec = np.asarray(Image.open(io.BytesIO(data)))
        with open(os.path.join(HERE, f"jp2_{name}.jp2"), "wb") as f:
            f.write(data)
        np.savez_compressed(os.path.join(HERE, f"jp2_{name}.npz"), source=src, openjpeg=dec)
        print(name, len(data))


if __name__ == "__main__":
    if os.path.isdir(REF):
        with open(os.path.join(HERE, "reference_tables.json"), "w") as f:
            json.dump(tables(), f)
    if "--package-only" not in sys.argv and "--jp2-only" not in sys.argv:
        codestreams()
    if "--jp2-only" not in sys.argv:
        package_streams()
    jp2_files()
