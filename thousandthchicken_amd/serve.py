"""Open-loop serving driver: Poisson arrivals of decode requests against a Server.

Requests are submitted at exponentially distributed gaps (rate = offered load
in requests per second) from the calling thread; a collector thread waits for
each request in submission order.  Latency is taken from the server's own clock
(submit to completion of the batch's result copy), so the driver's Python
overhead is not counted in it.  Throughput is decoded pixels over the span from
the first submission to the last completion.
"""
import threading
import time

import numpy as np


def arrivals(rate, n, seed):
    """n arrival times (seconds from 0) of a Poisson stream of the given rate."""
    g = np.random.default_rng(seed)
    return np.cumsum(g.exponential(1.0 / rate, size=n))


def mix(pool_small, pool_large, n, frac_large, seed):
    """Request sequence: indices into small (>= 0) or large (encoded as -1 - j) pools."""
    g = np.random.default_rng(seed + 1)
    out = []
    for _ in range(n):
        if pool_large and g.random() < frac_large:
            out.append(-1 - int(g.integers(len(pool_large))))
        else:
            out.append(int(g.integers(len(pool_small))))
    return out


def run(server, streams, rate, seed=0, keep=None):
    """Offers `streams` (list of codestreams, in order) at `rate` req/s.

    keep: optional set of positions whose decoded images are returned.
    Returns a dict of throughput and latency percentiles (and the kept images)."""
    n = len(streams)
    t_arr = arrivals(rate, n, seed)
    ids = [None] * n
    infos = [None] * n
    kept = {}
    ready = threading.Semaphore(0)
    err = []

    def collect():
        try:
            for i in range(n):
                ready.acquire()
                rid = ids[i]
                while not server.wait(rid, 5.0):
                    pass
                if keep is not None and i in keep:
                    img, info = server.result(rid)
                    kept[i] = img
                else:
                    info = server.info(rid)
                    server.release(rid)
                infos[i] = info
        except Exception as e:  # surfaced to the caller below
            err.append(e)

    th = threading.Thread(target=collect, daemon=True)
    th.start()
    t0 = time.perf_counter()
    t_sub = 0.0
    for i, cs in enumerate(streams):
        delay = t_arr[i] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ts = time.perf_counter()
        ids[i] = server.submit(cs)
        t_sub += time.perf_counter() - ts
        ready.release()
    th.join()
    if err:
        raise err[0]
    lat = np.array([x.done_s - x.submit_s for x in infos])
    wait = np.array([x.start_s - x.submit_s for x in infos])
    px = sum(x.width * x.height for x in infos)
    span = max(x.done_s for x in infos) - min(x.submit_s for x in infos)
    bs = np.array([x.batch_size for x in infos])
    failed = sum(1 for x in infos if x.status != 0)
    # latency of the second and the last fifth of the requests (in arrival
    # order): a server that keeps up holds it flat, a saturated one's queue
    # makes it grow with time
    q = max(1, n // 5)
    lat_early = float(np.median(lat[q:2 * q])) if n >= 2 * q else float(np.median(lat))
    lat_late = float(np.median(lat[n - q:]))
    stt = server.stats()
    return {
        "server_s": {"busy": round(stt.busy_s, 3), "plan": round(stt.plan_s, 3), "device": round(stt.device_s, 3),
                     "copy": round(stt.copy_s, 3), "batches": int(stt.batches), "submit": round(t_sub, 3)},
        "offered_rps": rate,
        "requests": n,
        "failed": failed,
        "achieved_rps": round(n / span, 2),
        "mpixel_s": round(px / span / 1e6, 2),
        "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
        "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
        "p50_queue_ms": round(float(np.percentile(wait, 50)) * 1e3, 3),
        "p50_ms_second_fifth": round(lat_early * 1e3, 3),
        "p50_ms_last_fifth": round(lat_late * 1e3, 3),
        "mean_batch": round(float(bs.mean()), 2),
        "span_s": round(span, 3),
        "images": kept,
    }


def serial(decoder, streams):
    """Per-image baseline: each request decoded alone (parse, upload, decode,
    download) one after the other on the same GPU.  Returns Mpixel/s and
    per-request latency percentiles."""
    lat = []
    px = 0
    t0 = time.perf_counter()
    for cs in streams:
        t = time.perf_counter()
        out = decoder.decode(cs)
        lat.append(time.perf_counter() - t)
        px += out.shape[0] * out.shape[1]
    span = time.perf_counter() - t0
    lat = np.array(lat)
    return {"requests": len(streams), "mpixel_s": round(px / span / 1e6, 2),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3), "span_s": round(span, 3)}
