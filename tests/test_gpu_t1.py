"""Every tier-1 kernel (one wave per block in stripe-column and row-mask form,
64 or 16 blocks per wave in lockstep) through its typed C-ABI launch on random
code-blocks of every shape, orientation, bit depth and truncation point,
against the oracle's tier-1 decode."""
import ctypes

import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle
from thousandthchicken_amd import synth

pytestmark = pytest.mark.gpu


def _blocks(seed, n, maxw=64, maxh=64):
    rng = np.random.default_rng(seed)
    L = synth._enc_lib()
    out = []
    for _ in range(n):
        w, h = int(rng.integers(1, maxw + 1)), int(rng.integers(1, maxh + 1))
        orient = int(rng.integers(0, 4))
        bits = int(rng.integers(1, 18))
        a = rng.integers(-(1 << bits) + 1, 1 << bits, (h, w)).astype(np.int32)
        a[rng.uniform(size=(h, w)) < rng.uniform(0, 0.95)] = 0
        buf = np.zeros(1 << 18, np.uint8)
        nb = ctypes.c_int(0)
        keep = int(rng.integers(1, 60)) if rng.uniform() < 0.5 else 999
        n_ = L.tce_t1_encode_truncated(a.ctypes.data, w, h, orient, keep, buf.ctypes.data, buf.size, ctypes.byref(nb))
        if nb.value == 0:
            continue
        npass = min(keep, 3 * nb.value - 2)
        out.append(dict(w=w, h=h, orient=orient, numbps=nb.value, npasses=npass, data=bytes(buf[:n_]), src=a))
    return out


def _expected(b):
    mag, sign, low, _ = oracle.t1_decode(b["data"], b["w"], b["h"], b["orient"], b["numbps"], b["npasses"])
    half = np.where(low >= 1, np.left_shift(1, np.maximum(low.astype(np.int32) - 1, 0)), 0)
    v = np.where(low >= 0, mag + half, 0)
    return np.where(sign.astype(bool), -v, v).astype(np.int32)


def _decode(kernel, small, blocks):
    """Decodes `blocks` with one tier-1 kernel through its C-ABI launch;
    returns the 64x64 output window of each block."""
    dev = tca.default_device()
    L = tca.lib()
    stream = ctypes.c_void_p()
    tca.check(L.tc_stream_create(dev.handle, ctypes.byref(stream)), "stream")
    arena = bytearray()
    descs = np.zeros(len(blocks), tca.CBLK_DESC)
    pitch = 64
    for i, b in enumerate(blocks):
        # the lockstep kernel takes codewords at any byte offset
        arena += bytes(i % 4) if kernel == "lockstep_unaligned" else bytes((-len(arena)) % 4)
        descs[i] = (len(arena), len(b["data"]), i * pitch * 64, pitch, b["w"], b["h"], b["numbps"], b["npasses"],
                    1 if b["orient"] == 1 else (2 if b["orient"] == 3 else 0), 0, 1.0)
        arena += b["data"]
    arena += bytes((-len(arena)) % 16) + b"\xff" * 64
    arena = np.frombuffer(bytes(arena), np.uint8)
    n_out = len(blocks) * pitch * 64
    max_passes = max(b["npasses"] for b in blocks)
    lbpw = 16 if kernel == "lockstep16" else 64
    scratch = L.tc_t1_lockstep_scratch_bytes(1 if small else 0, lbpw, len(blocks), max_passes)
    bufs = [tca.DeviceBuffer(dev, max(x, 256)) for x in (arena.nbytes, descs.nbytes, n_out * 4, scratch)]
    try:
        bufs[0].upload(stream, arena)
        bufs[1].upload(stream, descs)
        tca.check(L.tc_buf_fill32(stream, bufs[2].handle, 0, 0, n_out), "fill")
        if kernel == "stripe":
            st = L.tc_launch_t1_decode_stripe(stream, 1 if small else 0, 512, bufs[1].ptr, None, len(blocks),
                                              bufs[0].ptr, arena.nbytes, bufs[2].ptr)
        elif kernel in ("lockstep", "lockstep_unaligned", "lockstep16"):
            # the scratch state needs no initialisation: start it dirty
            tca.check(L.tc_buf_fill32(stream, bufs[3].handle, 0, 0xFFFFFFFF, scratch // 4), "fill")
            st = L.tc_launch_t1_decode_lockstep(stream, 1 if small else 0, lbpw, bufs[1].ptr, None, len(blocks),
                                                max_passes, bufs[0].ptr, arena.nbytes, bufs[3].ptr, scratch,
                                                bufs[2].ptr)
        else:
            st = L.tc_launch_t1_decode_wave(stream, 1 if small else 0, 512, bufs[1].ptr, None, len(blocks),
                                            bufs[0].ptr, arena.nbytes, bufs[2].ptr)
        tca.check(st, "launch")
        return bufs[2].download(stream, np.int32, n_out).reshape(len(blocks), 64, pitch)
    finally:
        for x in bufs:
            x.free()
        L.tc_stream_destroy(stream)


@pytest.mark.parametrize("kernel", ["stripe", "wave", "lockstep", "lockstep_unaligned", "lockstep16"])
@pytest.mark.parametrize("small", [True, False])
def test_t1_kernels_on_random_blocks(kernel, small):
    lim = 32 if small else 64
    blocks = _blocks(7 if small else 8, 300, lim, lim)
    print(f"t1 kernel={kernel} small={small} blocks={len(blocks)}", flush=True)
    out = _decode(kernel, small, blocks)
    for i, b in enumerate(blocks):
        got = out[i, :b["h"], :b["w"]]
        exp = _expected(b)
        assert np.array_equal(got, exp), (i, b["w"], b["h"], b["orient"], b["numbps"], b["npasses"])
        if b["npasses"] == 3 * b["numbps"] - 2:
            assert np.array_equal(got, b["src"])


def _byte_blocks(seed, n, lim):
    """Code-blocks whose codewords are byte strings rather than encoder
    output, weighted towards what the MQ byte-in treats specially: 0xFF bytes,
    0x80-0x8F after 0xFF (a carry into the stuffed bit), markers (above 0x8F
    after 0xFF) and codewords that end early."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        w, h = int(rng.integers(1, lim + 1)), int(rng.integers(1, lim + 1))
        numbps = int(rng.integers(1, 18))
        npasses = int(rng.integers(1, 3 * numbps - 1))
        n_bytes = int(rng.integers(0, 400))
        d = rng.integers(0, 256, n_bytes).astype(np.uint8)
        mode = i % 4
        for j in range(1, n_bytes):
            if mode >= 1 and rng.uniform() < (0.15 if mode == 1 else 0.3):
                d[j] = 0xFF
            if d[j - 1] == 0xFF:
                if mode == 2 or (mode == 3 and rng.uniform() < 0.7):
                    d[j] = 0x80 + int(rng.integers(0, 16))
                elif mode == 1 and d[j] > 0x8F:
                    d[j] &= 0x7F
        out.append(dict(w=w, h=h, orient=int(rng.integers(0, 4)), numbps=numbps, npasses=npasses,
                        data=d.tobytes()))
    return out


@pytest.mark.parametrize("kernel", ["stripe", "wave", "lockstep", "lockstep_unaligned", "lockstep16"])
@pytest.mark.parametrize("small", [True, False])
def test_t1_kernels_on_byte_strings(kernel, small):
    """Every kernel's MQ byte-in against T.800's on arbitrary codeword bytes:
    the decoded coefficients equal the oracle's bit for bit."""
    lim = 32 if small else 64
    blocks = _byte_blocks(11 if small else 12, 160, lim)
    print(f"t1 bytes kernel={kernel} small={small} blocks={len(blocks)}", flush=True)
    out = _decode(kernel, small, blocks)
    for i, b in enumerate(blocks):
        assert np.array_equal(out[i, :b["h"], :b["w"]], _expected(b)), (i, b["w"], b["h"], b["numbps"],
                                                                         b["npasses"], len(b["data"]))


def _edge_blocks(seed):
    """Dense and sparse blocks of edge shapes (single rows and columns, heights
    that are not a multiple of the 4-row stripe) and of up to 19 magnitude
    bits, full and truncated."""
    rng = np.random.default_rng(seed)
    L = synth._enc_lib()
    out = []
    shapes = [(1, 1), (1, 64), (64, 1), (2, 3), (3, 2), (5, 7), (64, 3), (3, 64), (63, 61), (64, 64)]
    for w, h in shapes:
        for bits, zero_frac, keep in ((19, 0.0, 999), (18, 0.5, 999), (12, 0.9, 999), (19, 0.0, 20), (8, 0.3, 7)):
            a = rng.integers(-(1 << bits) + 1, 1 << bits, (h, w)).astype(np.int32)
            a[rng.uniform(size=(h, w)) < zero_frac] = 0
            orient = int(rng.integers(0, 4))
            buf = np.zeros(1 << 20, np.uint8)
            nb = ctypes.c_int(0)
            n_ = L.tce_t1_encode_truncated(a.ctypes.data, w, h, orient, keep, buf.ctypes.data, buf.size,
                                           ctypes.byref(nb))
            if nb.value == 0:
                continue
            out.append(dict(w=w, h=h, orient=orient, numbps=nb.value, npasses=min(keep, 3 * nb.value - 2),
                            data=bytes(buf[:n_]), src=a))
    return out


def test_stripe_kernel_edge_shapes_and_deep_blocks():
    blocks = _edge_blocks(90)
    assert max(b["numbps"] for b in blocks) >= 19
    print(f"stripe edge blocks={len(blocks)}", flush=True)
    dev = tca.default_device()
    L = tca.lib()
    stream = ctypes.c_void_p()
    tca.check(L.tc_stream_create(dev.handle, ctypes.byref(stream)), "stream")
    arena = bytearray()
    descs = np.zeros(len(blocks), tca.CBLK_DESC)
    pitch = 64
    for i, b in enumerate(blocks):
        arena += bytes((-len(arena)) % 4)
        descs[i] = (len(arena), len(b["data"]), i * pitch * 64, pitch, b["w"], b["h"], b["numbps"], b["npasses"],
                    1 if b["orient"] == 1 else (2 if b["orient"] == 3 else 0), 0, 1.0)
        arena += b["data"]
    arena += bytes((-len(arena)) % 16) + b"\xff" * 64
    arena = np.frombuffer(bytes(arena), np.uint8)
    n_out = len(blocks) * pitch * 64
    bufs = [tca.DeviceBuffer(dev, max(x, 256)) for x in (arena.nbytes, descs.nbytes, n_out * 4)]
    try:
        bufs[0].upload(stream, arena)
        bufs[1].upload(stream, descs)
        tca.check(L.tc_buf_fill32(stream, bufs[2].handle, 0, 0, n_out), "fill")
        tca.check(L.tc_launch_t1_decode_stripe(stream, 0, 64, bufs[1].ptr, None, len(blocks), bufs[0].ptr,
                                               arena.nbytes, bufs[2].ptr), "launch")
        out = bufs[2].download(stream, np.int32, n_out).reshape(len(blocks), 64, pitch)
        for i, b in enumerate(blocks):
            got = out[i, :b["h"], :b["w"]]
            assert np.array_equal(got, _expected(b)), (i, b["w"], b["h"], b["orient"], b["numbps"], b["npasses"])
            if b["npasses"] == 3 * b["numbps"] - 2:
                assert np.array_equal(got, b["src"])
    finally:
        for x in bufs:
            x.free()
        L.tc_stream_destroy(stream)
