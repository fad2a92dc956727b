"""Code-block styles (BYPASS, RESET, TERMALL, CAUSAL, PTERM, SEGSYM) on the GPU: the lockstep
tier-1 kernel through its C-ABI launch on random styled code-blocks against the
oracle's styled tier-1, and whole styled codestreams through the fused, batched
and staged paths against the oracle (lossless bit-exact, lossy within 1)."""
import ctypes

import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle
from tests.test_styles import STYLES, reconstruct, styled_blocks
from thousandthchicken_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bpw", [64, 16])
@pytest.mark.parametrize("small", [True, False])
@pytest.mark.parametrize("style", STYLES)
def test_lockstep_on_styled_blocks(style, small, bpw):
    lim = 32 if small else 64
    blocks = styled_blocks(500 + style + small, 200, style, lim, lim)
    print(f"styled t1 style={style} small={small} bpw={bpw} blocks={len(blocks)}", flush=True)
    dev = tca.default_device()
    L = tca.lib()
    stream = ctypes.c_void_p()
    tca.check(L.tc_stream_create(dev.handle, ctypes.byref(stream)), "stream")
    mode = {2: 2, 8: 4, 32: 8, 1: 16, 4: 32}  # style bit -> descriptor mode bit
    mbits = sum(v for k, v in mode.items() if style & k)
    arena = bytearray()
    descs = np.zeros(len(blocks), tca.CBLK_DESC)
    pitch = 64
    for i, b in enumerate(blocks):
        if style & 5:
            # segment lengths and their count in the words before the codeword
            arena += bytes((-len(arena)) % 4)
            arena += np.asarray(b["seglen"] + [len(b["seglen"])], "<u4").tobytes()
        else:
            arena += bytes(i % 3)
        descs[i] = (len(arena), len(b["data"]), i * pitch * 64, pitch, b["w"], b["h"], b["numbps"], b["npasses"],
                    1 if b["orient"] == 1 else (2 if b["orient"] == 3 else 0), mbits, 1.0)
        arena += b["data"]
    arena += bytes((-len(arena)) % 16) + b"\xff" * 64
    arena = np.frombuffer(bytes(arena), np.uint8)
    n_out = len(blocks) * pitch * 64
    max_passes = max(b["npasses"] for b in blocks)
    scratch = L.tc_t1_lockstep_scratch_bytes(1 if small else 0, bpw, len(blocks), max_passes)
    bufs = [tca.DeviceBuffer(dev, max(x, 256)) for x in (arena.nbytes, descs.nbytes, n_out * 4, scratch)]
    try:
        bufs[0].upload(stream, arena)
        bufs[1].upload(stream, descs)
        tca.check(L.tc_buf_fill32(stream, bufs[2].handle, 0, 0, n_out), "fill")
        tca.check(L.tc_buf_fill32(stream, bufs[3].handle, 0, 0xFFFFFFFF, scratch // 4), "fill")
        tca.check(L.tc_launch_t1_decode_lockstep(stream, 1 if small else 0, bpw, bufs[1].ptr, None, len(blocks),
                                                 max_passes, bufs[0].ptr, arena.nbytes, bufs[3].ptr, scratch,
                                                 bufs[2].ptr), "launch")
        out = bufs[2].download(stream, np.int32, n_out).reshape(len(blocks), 64, pitch)
        for i, b in enumerate(blocks):
            got = out[i, :b["h"], :b["w"]]
            exp = reconstruct(*oracle.t1_decode(b["data"], b["w"], b["h"], b["orient"], b["numbps"], b["npasses"],
                                                style, b["seglen"])[:3])
            assert np.array_equal(got, exp), (i, b["w"], b["h"], b["orient"], b["numbps"], b["npasses"])
            if b["npasses"] == 3 * b["numbps"] - 2:
                assert np.array_equal(got, b["src"])
    finally:
        for x in bufs:
            x.free()
        L.tc_stream_destroy(stream)


def _streams():
    out = []
    for i, style in enumerate(STYLES):
        src = synth.image(70 + 9 * i, 90 + 13 * i, 3 if i % 2 == 0 else 1, 40 + i)
        lossy = i % 3 == 2
        cb = 32 if i % 2 else 64
        out.append((style, lossy, synth.encode(src, levels=4, cb=cb, cbsty=style, lossy=lossy, bpp=2.5), src))
    return out


def _check(out, cs, lossy, src):
    ref = oracle.decode(cs).transpose(1, 2, 0)
    d = np.abs(out.reshape(ref.shape).astype(np.int32) - ref)
    assert d.max() <= (1 if lossy else 0)
    if not lossy:
        assert np.array_equal(out.reshape(src.shape), src)


@pytest.mark.parametrize("case", range(len(STYLES)))
def test_styled_codestreams_fused_and_staged(gpu_decoder, case):
    style, lossy, cs, src = _streams()[case]
    print(f"styled image style={style} lossy={lossy} {src.shape}", flush=True)
    _check(gpu_decoder.decode(cs), cs, lossy, src)
    _check(tca.decode_staged(cs), cs, lossy, src)


@pytest.mark.parametrize("bpw", [0, 1, tca.BPW_LOCKSTEP32])
def test_styled_and_plain_codestreams_in_one_batch(bpw):
    cases = _streams()
    plain = synth.image(100, 120, 3, 77)
    cases.append((0, False, synth.encode(plain, levels=5, cb=32), plain))
    print(f"styled batch of {len(cases)} bpw={bpw}", flush=True)
    dec = tca.Decoder()
    try:
        dec.set_blocks_per_wave(True, bpw)
        dec.set_blocks_per_wave(False, bpw)
        dec.prepare([c[2] for c in cases])
        dec.run()
        dec.sync()
        for i, (style, lossy, cs, src) in enumerate(cases):
            _check(dec.fetch(i), cs, lossy, src)
    finally:
        dec.close()
