"""The boundary as a maintainer of the reference would use it: the C++ example
of INTEGRATION.md section 2, extracted from the document as written, is built
with g++ against include/tc/*.h and libtc.so (on the CPU: it must compile and
link, so every class and C entry point it names is declared and exported), then
run on a golden codestream (on the GPU: its output file must equal the source
image).  tc_buf_copy_rect, the replacement of clEnqueueCopyBufferRect
(TC/DWTKernel.cpp:135-147, TC/Quantizer.cpp:72-115), is checked against a numpy
slice of the same bytes."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import thousandthchicken_amd as tca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def snippet():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, re.S)
    assert blocks, "INTEGRATION.md holds no C++ example"
    return blocks[0]


def build_snippet(tmp):
    src = os.path.join(tmp, "snippet.cpp")
    with open(src, "w") as f:
        f.write(snippet())
    exe = os.path.join(tmp, "snippet")
    libdir = os.path.join(ROOT, "thousandthchicken_amd")
    cmd = ["g++", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"), src, "-L" + libdir, "-ltc",
           "-Wl,-rpath," + libdir, "-Wl,-rpath-link,/opt/rocm/lib", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_integration_snippet_compiles_and_links(tmp_path):
    assert os.path.exists(build_snippet(str(tmp_path)))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["openjpeg_ll_rgb64", "openjpeg_ll_gray16", "openjpeg_ll_tiles_offset"])
def test_integration_snippet_decodes_a_codestream(tmp_path, name):
    exe = build_snippet(str(tmp_path))
    out = os.path.join(str(tmp_path), "out.raw")
    print(f"snippet on {name}", flush=True)
    r = subprocess.run([exe, os.path.join(GOLDEN, name + ".j2k"), out], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    src = np.load(os.path.join(GOLDEN, name + ".npz"))["source"]
    got = np.fromfile(out, src.dtype)
    assert np.array_equal(got, src.ravel())


@pytest.mark.gpu
def test_buf_copy_rect_equals_numpy_slice():
    L = tca.lib()
    P = ctypes.POINTER
    L.tc_buf_copy_rect.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, P(ctypes.c_size_t),
                                   P(ctypes.c_size_t), P(ctypes.c_size_t)] + [ctypes.c_size_t] * 4
    dev = tca.default_device()
    stream = ctypes.c_void_p()
    tca.check(L.tc_stream_create(dev.handle, ctypes.byref(stream)), "stream")
    rng = np.random.default_rng(3)
    # source: 3 slices of 37 rows of 211 bytes, row pitch 224, slice pitch 224 * 40
    srp, ssp = 224, 224 * 40
    src = rng.integers(0, 256, 3 * ssp, dtype=np.uint8)
    drp, dsp = 160, 160 * 20
    a, b = tca.DeviceBuffer(dev, src.nbytes), tca.DeviceBuffer(dev, 3 * dsp)
    try:
        a.upload(stream, src)
        b.upload(stream, np.zeros(3 * dsp, np.uint8))
        so = (ctypes.c_size_t * 3)(5, 2, 1)      # byte x, row, slice
        do = (ctypes.c_size_t * 3)(3, 4, 0)
        region = (ctypes.c_size_t * 3)(150, 13, 2)  # bytes per row, rows, slices
        tca.check(L.tc_buf_copy_rect(stream, b.handle, a.handle, so, do, region, srp, ssp, drp, dsp), "copy_rect")
        got = b.download(stream, np.uint8, 3 * dsp)
        exp = np.zeros(3 * dsp, np.uint8)
        s3 = src.reshape(3, 40, 224)
        e3 = exp.reshape(3, 20, 160)
        e3[0:2, 4:17, 3:153] = s3[1:3, 2:15, 5:155]
        assert np.array_equal(got, exp)
        # a region reaching past the source (its second slice) fails and copies nothing
        b.upload(stream, np.zeros(3 * dsp, np.uint8))
        bad = (ctypes.c_size_t * 3)(150, 40, 2)
        assert L.tc_buf_copy_rect(stream, b.handle, a.handle, so, do, bad, srp, ssp, drp, dsp) != 0
        assert not b.download(stream, np.uint8, 3 * dsp).any()
    finally:
        a.free()
        b.free()
        L.tc_stream_destroy(stream)
