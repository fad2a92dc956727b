"""Decode server (continuous batching): every result equals the same request
decoded alone, admission respects the per-slot memory share, a broken request
fails alone, and an open-loop Poisson run completes every request."""
import numpy as np
import pytest

import thousandthchicken_amd as tca
from oracle import oracle
from thousandthchicken_amd import serve, synth, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def streams():
    s = workloads.cfg4(n=20, scale=0.12)
    big = synth.encode(synth.image(300, 520, 1, 77, 16), bits=16, levels=6, cb=64)
    return s + [big]


def _alone(streams):
    srv = tca.Server(slots=1, max_batch=1)
    out = []
    for cs in streams:
        rid = srv.submit(cs)
        assert srv.wait(rid, 60.0)
        img, info = srv.result(rid)
        assert info.batch_size == 1
        out.append(img)
    srv.close()
    return out


def test_alone_equals_mix(streams):
    print(f"server alone vs mix, {len(streams)} requests", flush=True)
    alone = _alone(streams)
    srv = tca.Server(slots=3, max_batch=64)
    ids = [srv.submit(cs) for cs in streams]
    infos = []
    for i, rid in enumerate(ids):
        assert srv.wait(rid, 60.0)
        img, info = srv.result(rid)
        infos.append(info)
        assert img.dtype == alone[i].dtype and img.shape == alone[i].shape
        assert np.array_equal(img, alone[i]), i
    assert max(x.batch_size for x in infos) > 1, "requests were never batched together"
    st = srv.stats()
    assert st.completed == len(streams) and st.failed == 0
    srv.close()
    # and both agree with the oracle (lossless exact, lossy within 1 LSB)
    lay = workloads.cfg4_layout(20)
    for i, cs in enumerate(streams):
        ref = oracle.decode(cs).transpose(1, 2, 0)
        lim = 0 if (i >= 20 or lay[i]["lossless"]) else 1
        assert np.abs(alone[i].astype(np.int32) - ref).max() <= lim, i


def test_admission_by_footprint(streams):
    small = streams[:12]
    fp = []
    for cs in small:
        srv = tca.Server(slots=1, max_batch=1)
        rid = srv.submit(cs)
        assert srv.wait(rid, 60.0)
        fp.append(srv.info(rid).footprint)
        srv.release(rid)
        srv.close()
    share = 2 * max(fp)
    srv = tca.Server(slots=2, max_batch=256, memory_budget=2 * share)
    ids = [srv.submit(cs) for cs in small]
    batches = {}
    for rid in ids:
        assert srv.wait(rid, 60.0)
        info = srv.info(rid)
        batches.setdefault(info.batch, []).append(info.footprint)
        srv.release(rid)
    for b, f in batches.items():
        assert len(f) == 1 or sum(f) <= share, (b, f)
    assert srv.stats().peak_bytes <= 2 * share
    srv.close()


@pytest.mark.parametrize("bpw", [0, tca.BPW_LOCKSTEP16])
def test_slot_device_bytes_within_share(bpw):
    # requests of mixed sizes; under lockstep-16 packing every block holds
    # tier-1 state in the scratch buffer, the largest share of the footprint
    reqs = workloads.cfg4(n=24, scale=0.35) + [synth.encode(synth.image(700, 900, 1, 78, 16), bits=16, levels=5,
                                                            cb=64)]
    fp = []
    for cs in reqs:
        # alone, the resident buffers of the slot stay within the request's
        # estimated footprint, grown by the slot's 1.25 factor (1 MiB floor per buffer)
        srv = tca.Server(slots=1, max_batch=1, bpw_small=bpw, bpw_large=bpw)
        rid = srv.submit(cs)
        assert srv.wait(rid, 60.0)
        f = srv.info(rid).footprint
        fp.append(f)
        srv.release(rid)
        held = srv.stats().slot_device_peak
        srv.close()
        assert 0 < held <= 1.25 * f + (4 << 20), (held, f)
    share = 3 * max(fp)
    print(f"slot share {share >> 20} MiB, footprints {min(fp) >> 20}..{max(fp) >> 20} MiB", flush=True)
    srv = tca.Server(slots=2, max_batch=256, memory_budget=2 * share, bpw_small=bpw, bpw_large=bpw)
    ids = [srv.submit(cs) for cs in reqs * 2]
    sizes = []
    for rid in ids:
        assert srv.wait(rid, 60.0)
        sizes.append(srv.info(rid).batch_size)
        srv.release(rid)
    st = srv.stats()
    srv.close()
    assert max(sizes) > 1
    assert st.slot_device_peak <= share, (st.slot_device_peak, share)


def test_bad_request_fails_alone(streams):
    srv = tca.Server(slots=2)
    good = srv.submit(streams[0])
    bad = srv.submit(b"\xff\x4f\xff\x51" + b"\x00" * 40)
    assert srv.wait(good, 60.0) and srv.wait(bad, 60.0)
    assert srv.info(bad).status != 0
    with pytest.raises(tca.TCError):
        srv.result(bad)
    img, info = srv.result(good)
    assert info.status == 0 and img.size > 0
    srv.close()


def test_broken_packets_fail_alone_in_batch(streams):
    # main header intact (accepted at submit), tile-part marker broken (fails in the worker's parse)
    bad = bytearray(streams[1])
    k = bytes(bad).find(b"\xff\x90")
    assert k > 0
    bad[k + 1] = 0x00
    srv = tca.Server(slots=1, max_batch=64)
    ids = [srv.submit(streams[0]), srv.submit(bytes(bad)), srv.submit(streams[2])]
    for rid in ids:
        assert srv.wait(rid, 60.0)
    assert srv.info(ids[1]).status != 0
    for rid, cs in ((ids[0], streams[0]), (ids[2], streams[2])):
        img, info = srv.result(rid)
        ref = oracle.decode(cs).transpose(1, 2, 0)
        assert np.abs(img.astype(np.int32) - ref).max() <= 1
    srv.release(ids[1])
    assert srv.stats().failed == 1
    srv.close()


def test_poisson_run(streams):
    srv = tca.Server(slots=3)
    order = serve.mix(streams[:20], streams[20:], 60, 0.1, seed=9)
    seq = [streams[i] if i >= 0 else streams[20 + (-1 - i)] for i in order]
    r = serve.run(srv, seq, rate=400.0, seed=9, keep={0, 5, 17})
    assert r["failed"] == 0 and r["requests"] == 60
    assert r["p99_ms"] >= r["p50_ms"] > 0
    for i, img in r["images"].items():
        ref = oracle.decode(seq[i]).transpose(1, 2, 0)
        assert np.abs(img.astype(np.int32) - ref).max() <= 1
    srv.close()
