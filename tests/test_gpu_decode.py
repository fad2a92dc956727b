"""GPU decode parity: the HIP path through the C ABI against the CPU oracle.

Lossless (5/3) must be bit-exact with the oracle and with the source image;
lossy (9/7) must be within 1 LSB of the oracle's FP64 decode.
"""
import numpy as np
import pytest

from oracle import oracle
from tests import imaging
from thousandthchicken_amd import synth

pytestmark = pytest.mark.gpu

LOSSLESS = [
    ("rgb64_l2", (64, 64, 3), dict(num_resolutions=3, mct=1)),
    ("rgb100x77_cb32", (100, 77, 3), dict(num_resolutions=4, codeblock_size=(32, 32), mct=1)),
    ("rgb512_cfg1", (512, 512, 3), dict(num_resolutions=6, codeblock_size=(64, 64), mct=1)),
    ("rgb_tiles_off", (300, 200, 3), dict(num_resolutions=4, tile_size=(64, 64), offset=(3, 5), mct=1)),
    ("gray_nomct", (130, 257, 1), dict(num_resolutions=5)),
    ("rgb_nomct", (96, 80, 3), dict(num_resolutions=3, mct=0)),
    ("rgb_l0", (40, 30, 3), dict(num_resolutions=1, mct=1)),
    ("rgb_prec_pcrl", (300, 200, 3), dict(num_resolutions=5, codeblock_size=(32, 32), precinct_size=(128, 128),
                                          progression="PCRL", mct=1)),
    ("rgb_prec_cprl_tiles", (257, 300, 3), dict(num_resolutions=4, codeblock_size=(16, 16), precinct_size=(64, 64),
                                                progression="CPRL", tile_size=(128, 128), offset=(5, 3), mct=1)),
]

LOSSY = [
    ("lossy256", (256, 256, 3), dict(quality_layers=[24], num_resolutions=6, mct=1)),
    ("lossy_odd", (255, 301, 3), dict(quality_layers=[12], num_resolutions=5, mct=1)),
    ("lossy_layers", (200, 220, 3), dict(quality_layers=[40, 20, 10], num_resolutions=5, mct=1)),
    ("lossy_gray", (190, 120, 1), dict(quality_layers=[8], num_resolutions=4)),
    ("lossy_prec_rlcp", (240, 260, 3), dict(quality_layers=[20, 8], num_resolutions=5, codeblock_size=(32, 32),
                                            precinct_size=(128, 128), progression="RLCP", mct=1)),
]


def _img(shape, seed):
    h, w, c = shape
    a = synth.image(h, w, c, seed)
    return a[..., 0] if c == 1 else a


@pytest.mark.parametrize("name,shape,kw", LOSSLESS, ids=[c[0] for c in LOSSLESS])
def test_lossless_bit_exact(gpu_decoder, name, shape, kw):
    if imaging.pillow_j2k() is None:
        pytest.skip("OpenJPEG not available to make the codestream")
    src = _img(shape, 11)
    cs = imaging.encode_openjpeg(src, **kw)
    print(f"case {name} {shape} {len(cs)} bytes", flush=True)
    out = gpu_decoder.decode(cs)
    ref = oracle.decode(cs)
    assert np.array_equal(out.transpose(2, 0, 1).astype(np.int32), ref)
    assert np.array_equal(out.reshape(src.shape), src)


@pytest.mark.parametrize("name,shape,kw", LOSSY, ids=[c[0] for c in LOSSY])
def test_lossy_within_one(gpu_decoder, name, shape, kw):
    if imaging.pillow_j2k() is None:
        pytest.skip("OpenJPEG not available to make the codestream")
    src = _img(shape, 12)
    cs = imaging.encode_openjpeg(src, irreversible=True, quality_mode="rates", **kw)
    print(f"case {name} {shape} {len(cs)} bytes", flush=True)
    out = gpu_decoder.decode(cs).transpose(2, 0, 1).astype(np.int32)
    ref, reff = oracle.decode(cs, want_float=True)
    d = np.abs(out - ref)
    assert d.max() <= 1, f"max |gpu - oracle| = {d.max()}"
    assert (d > 0).mean() < 0.01


@pytest.mark.parametrize("mode,shape", [("cinema2k-24", (1080, 2048, 3)), ("cinema4k-24", (2160, 4096, 3))])
def test_cinema_profiles_within_one(gpu_decoder, mode, shape):
    """OpenJPEG's Digital Cinema profiles at full frame size: CPRL with
    precincts, tile-parts, TLM and (4K) a tile-header POC."""
    if imaging.pillow_j2k() is None:
        pytest.skip("OpenJPEG not available to make the codestream")
    src = _img(shape, 13)
    cs = imaging.encode_openjpeg(src, cinema_mode=mode)
    print(f"case {mode} {shape} {len(cs)} bytes", flush=True)
    out = gpu_decoder.decode(cs).transpose(2, 0, 1).astype(np.int32)
    ref = oracle.decode(cs)
    d = np.abs(out - ref)
    assert d.max() <= 1, f"max |gpu - oracle| = {d.max()}"
    assert (d > 0).mean() < 0.01


@pytest.mark.parametrize("packed", ["ppt", "ppm"])
def test_packed_headers_equal_inline(gpu_decoder, packed):
    """PPT / PPM packed packet headers (package encoder, several tiles and
    small segments): the GPU image equals the in-line stream's and the source."""
    import thousandthchicken_amd as tca
    src = _img((300, 260, 3), 14)
    inline = synth.encode(src, levels=4, cb=32, tile=(128, 128))
    cs = synth.encode(src, levels=4, cb=32, tile=(128, 128), packed=packed, packed_max=64)
    print(f"case {packed} {len(cs)} bytes", flush=True)
    out = gpu_decoder.decode(cs)
    assert np.array_equal(out, gpu_decoder.decode(inline))
    assert np.array_equal(out.reshape(src.shape), src)
    assert np.array_equal(tca.decode_staged(cs).reshape(src.shape), src)
