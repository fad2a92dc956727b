"""Decode throughput benchmark (decoded Mpixel/s on one MI355X).

Headline (`value`): the mixed batch of 256 images -- sides log-uniform in
[256, 2048] per axis, 1 or 3 components, 8-bit, half lossless (RCT + 5/3,
5 levels) and half lossy (ICT + 9/7 at 0.5-2 bpp), 32x32 or 64x64 code-blocks,
image seeds 1000-1255, coded by the package's encoder -- decoded with one
launch per stage over the whole batch and replayed from a captured graph,
codestreams resident in HBM.  A step decodes the whole batch: fill, tier-1
(one launch per code-block size class, the two side by side), the inverse-DWT
levels and the fused final level.  Steady state keeps several batches in
flight: copies of the prepared batch on their own pipelines and streams,
launched round robin, so that one batch's tier-1 tail (its longest lockstep
waves, which leave most compute units idle) overlaps the next batch's work;
the number in flight and the tier-1 packing are picked by measurement, and
one batch alone is reported beside it.  Outputs are checked against
per-image decode first, and every in-flight copy's outputs again after
timing (lossless identical, lossy within 1 LSB).

Beside it the line carries (unless --quick):
  per_image  the 512x512 RGB lossless image, the 2048x2048 RGB lossy image and
             the 7680x4320 16-bit lossless image through the per-image entry
             (host codestream in, host image out, blocking): Mpixel/s, p50 and
             p99 latency, and the decode time with the codestream resident;
  accuracy   the 2048x2048 lossy image and the batch's lossy images against the
             FP64 oracle: max |gpu - rounded FP64|, and max and RMS error
             against the unrounded FP64 samples;
  batched    the 256-image batch eager and from the graph against the same
             images decoded one by one, and the 3840x2160 12-bit image (twelve
             1024x1024 tiles) batched against per image;
  served     Poisson arrivals (90% batch images, 10% 7680x4320 16-bit) at
             100..5000 req/s, each load for at least 5 s or 3000 requests,
             offered against achieved rate, p50/p99 latency, against serial
             per-image decode of the same mix;
  reduced, forward, region: the other decode forms;
  cpu_baseline: the C oracle on a bounded sample of the batch's images, on all
             host cores (one image per thread) and on one core.

Prints one JSON line.
"""
import argparse
import concurrent.futures as cf
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pct(a, q):
    return round(float(np.percentile(np.asarray(a), q)) * 1e3, 3)


def cpu_threads():
    # the GPU box gives one GPU a share of 16 host cores
    return max(1, min(16, os.cpu_count() or 1))


def cpu_baseline(streams, budget_s=8.0):
    """The oracle (scalar C) decoding the batch's images in order: on one core
    until budget_s has passed, then the same images on all host cores (one
    image per thread; the oracle's C calls release the GIL)."""
    from oracle import oracle

    def px_of(cs):
        a = oracle.decode(cs)
        return a.shape[1] * a.shape[2]
    t0 = time.perf_counter()
    n = px = 0
    while time.perf_counter() - t0 < budget_s and n < len(streams):
        px += px_of(streams[n])
        n += 1
    one = px / (time.perf_counter() - t0) / 1e6
    sample = streams[:n]
    P = cpu_threads()
    reps = max(1, int(round(P * 0.75)))  # about budget_s again on P cores
    work = sample * reps
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(P) as ex:
        tot = sum(ex.map(px_of, work))
    allc = tot / (time.perf_counter() - t0) / 1e6
    return {"value": round(allc, 3), "unit": "Mpixel/s", "cores": P, "kind": "port",
            "sample": f"the first {n} images of the 256-image batch, {px / 1e6:.1f} Mpixel, decoded by the C oracle "
                      f"{reps} times over on {P} threads",
            "one_core": {"value": round(one, 3), "unit": "Mpixel/s", "cores": 1,
                         "sample": f"the same {n} images once on one core"}}


def lossy_errors(out_hwc, cs):
    """(max |gpu - rounded FP64|, sum of squared errors, max |error|, samples)
    of a lossy decode against the FP64 oracle's unrounded samples (clamped)."""
    from oracle import oracle
    import thousandthchicken_amd as tca
    ref, ref_f = oracle.decode(cs, want_float=True)
    ref = ref.transpose(1, 2, 0).astype(np.int32)
    hi = (1 << tca.probe(cs).max_prec) - 1
    err = out_hwc.reshape(ref.shape).astype(np.int32) - np.clip(ref_f.transpose(1, 2, 0), 0, hi)
    return int(np.abs(out_hwc.reshape(ref.shape).astype(np.int32) - ref).max()), float((err * err).sum()), \
        float(np.abs(err).max()), err.size


def error_summary(parts):
    n = sum(p[3] for p in parts)
    return {"max_lsb_vs_rounded_fp64": max(p[0] for p in parts),
            "max_abs_vs_fp64": round(max(p[2] for p in parts), 4),
            "rms_vs_fp64": round((sum(p[1] for p in parts) / n) ** 0.5, 4), "samples": n}


def check_equal_source(out, src, what):
    if not np.array_equal(out.reshape(src.shape), src):
        raise SystemExit(f"{what}: decode is not bit-exact")


def stage_times(dec, reps=5):
    """Mean stage times of timed runs, and each tier-1 launch's time on its own stream."""
    keys = ("fill_ms", "t1_ms", "dwt_ms", "final_ms", "total_ms")
    acc = {k: 0.0 for k in keys}
    launch = [0.0] * 4
    for _ in range(reps):
        dec.run(timed=True)
        dec.sync()
        s = dec.stats()
        for k in keys:
            acc[k] += getattr(s, k) / reps
        for g in range(s.t1_launches):
            launch[g] += s.t1_launch_ms[g] / reps
    return acc, launch


def timed_runs(dec, steps, graph=False):
    dec.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        dec.launch_graph() if graph else dec.run()
    dec.sync()
    return (time.perf_counter() - t0) / steps


def in_flight_leg(tca, streams, outs, lossless, px, args, tuned):
    """The batch decoded back to back with D batches in flight: D resident
    copies of the prepared batch, each its own pipeline (buffers, streams,
    captured graph), launched round robin.  A step is still one whole decode of
    the batch; the copies fill the compute units that one batch's longest
    tier-1 waves leave idle.  With more than one in flight, the tier-1 packing
    is picked for throughput (fewer wave-steps) rather than for one batch's
    latency.  Every copy's outputs are checked against per-image decode."""
    depths = sorted({int(x) for x in args.in_flight.split(",") if x and int(x) >= 1})
    cands = [(tuned["bpw_small"], tuned["bpw_large"])] if tuned else [(args.bpw_small, args.bpw_large)]
    if not args.no_autotune:
        cands += [(-64, -16), (-64, -32), (-64, -64)]
    cands = list(dict.fromkeys(cands))
    dmax = max(depths)
    tried, best = [], None
    for bs, bl in cands:
        decs = []
        for _ in range(dmax):
            d = tca.Decoder()
            d.set_blocks_per_wave(True, bs)
            d.set_blocks_per_wave(False, bl)
            if tuned and tuned.get("dwt_shape") is not None:
                t = d.tuning()
                t.dwt_shape = tuned["dwt_shape"]
                d.set_tuning(t)
            d.prepare(streams)
            d.capture_graph()
            decs.append(d)
        for depth in depths:
            for k in range(2 * depth):
                decs[k % depth].launch_graph()
            for d in decs:
                d.sync()
            t0 = time.perf_counter()
            for k in range(args.steps):
                decs[k % depth].launch_graph()
            for d in decs[:depth]:
                d.sync()
            ms = (time.perf_counter() - t0) / args.steps * 1e3
            r = {"bpw_small": bs, "bpw_large": bl, "in_flight": depth, "ms_per_batch": round(ms, 3),
                 "mpixel_s": round(px / ms / 1e3, 1)}
            tried.append(r)
            log(f"in flight: {r}")
            if best is None or ms < best["ms_per_batch"]:
                best = dict(r)
        for d in decs:
            d.close()
    # the chosen configuration again, every copy's outputs checked
    decs = []
    for _ in range(best["in_flight"]):
        d = tca.Decoder()
        d.set_blocks_per_wave(True, best["bpw_small"])
        d.set_blocks_per_wave(False, best["bpw_large"])
        if tuned and tuned.get("dwt_shape") is not None:
            t = d.tuning()
            t.dwt_shape = tuned["dwt_shape"]
            d.set_tuning(t)
        d.prepare(streams)
        d.capture_graph()
        decs.append(d)
    depth = len(decs)
    for k in range(args.warmup * depth):
        decs[k % depth].launch_graph()
    for d in decs:
        d.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        decs[k % depth].launch_graph()
    for d in decs:
        d.sync()
    sec = (time.perf_counter() - t0) / args.steps
    for j, d in enumerate(decs):
        for i in range(len(streams)):
            diff = int(np.abs(d.fetch(i).astype(np.int32) - outs[i].astype(np.int32)).max())
            if diff > (0 if lossless[i] else 1):
                raise SystemExit(f"in-flight copy {j}, image {i}: differs from per-image decode by {diff}")
    for d in decs:
        d.close()
    best["ms_per_batch"] = round(sec * 1e3, 3)
    best["mpixel_s"] = round(px / sec / 1e6, 1)
    return best, tried, sec


def t1_kernel_name(bpw, small, args):
    """The tier-1 kernel a launch runs, as rocprof names it."""
    if bpw < 0:
        return "t1_lockstep_kernel<unsigned int, 32>" if small else "t1_lockstep_kernel<unsigned long, 64>"
    if args.t1_kernel == 0:
        return "t1_stripe_kernel<unsigned int, 32, 32>" if small else "t1_stripe_kernel<unsigned long, 20, 64>"
    return "t1_wave_kernel"


def kernel_tag(name):
    """File-name form of a kernel name: t1_lockstep_kernel_64, t1_stripe_kernel_32, ..."""
    base = name.split("<")[0]
    if "<" not in name:
        return base
    return f"{base}_{name.rstrip('>').split(',')[-1].strip()}"


def t1_launches(s, ms, args):
    return [{"kernel": t1_kernel_name(s.t1_launch_bpw[g], s.t1_launch_class[g] == 0, args),
             "blocks": s.t1_launch_blocks[g], "blocks_per_wave": s.t1_launch_bpw[g],
             "bytes": s.t1_launch_bytes[g], "ms": round(ms[g], 3)} for g in range(s.t1_launches)]


def traffic_record(kernel, workload):
    """HBM bytes per launch measured with FETCH_SIZE / WRITE_SIZE
    (profiles/traffic_<kernel>_<workload>.json)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel_tag(kernel)}_{workload}.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def valu_record(workload):
    """Wave-wide VALU instructions of one batch's tier-1 launches at a given
    packing, counted with SQ_INSTS_VALU (profiles/valu_<workload>.json)."""
    try:
        with open(os.path.join(ROOT, "profiles", f"valu_{workload}.json")) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def roofline_valu(flight, workload):
    """Tier-1's VALU issue rate in the chosen steady state against the
    machine's, when the counted packing is the chosen one (else None).  The
    time per batch also holds the DWT stages, so this understates tier-1's own
    rate."""
    rec = valu_record(workload)
    if rec is None or flight["in_flight"] < 2 or rec.get("packing") != {"bpw_small": flight["bpw_small"],
                                                                        "bpw_large": flight["bpw_large"]}:
        return None
    achieved = rec["valu_tier1_per_batch"] / (flight["ms_per_batch"] * 1e-3)
    peak = rec["peak_wave_instructions_per_s"]
    return {"bound": "valu", "kernel": "t1_lockstep_kernel (both block classes)",
            "achieved": round(achieved / 1e9, 1), "peak": round(peak / 1e9, 1), "unit": "G wave-instructions/s",
            "frac": round(achieved / peak, 4), "valu_per_batch": rec["valu_tier1_per_batch"],
            "source": f"profiles/valu_{workload}.json",
            "note": "tier-1 VALU instructions of one batch (counted) over the measured time per batch in flight; "
                    "peak: 1024 SIMDs x 2.4 GHz, one wave-wide instruction per 4 cycles"}


def tuned_decoder(tca, args):
    dec = tca.Decoder()
    dec.set_blocks_per_wave(False, args.bpw_large)
    dec.set_blocks_per_wave(True, args.bpw_small)
    dec.set_t1_kernel(args.t1_kernel)
    if args.dwt_shape >= 0:
        t = dec.tuning()
        t.dwt_shape = args.dwt_shape
        dec.set_tuning(t)
    return dec


def per_image_leg(tca, cs, src, calls, what, resident_steps=10):
    """Per-image entry (host in, host out), one call after the other: latency
    percentiles and Mpixel/s; then the resident decode time of the same image."""
    dec = tca.Decoder()
    out = dec.decode(cs)
    if src is not None:
        check_equal_source(out, src, what)
    lat = []
    for _ in range(calls):
        t = time.perf_counter()
        dec.decode(cs)
        lat.append(time.perf_counter() - t)
    dec.prepare([cs])
    dec.run()
    res_s = timed_runs(dec, resident_steps)
    dec.close()
    i = tca.probe(cs)
    px = i.width * i.height
    return {"width": i.width, "height": i.height, "components": i.ncomp, "bits": i.max_prec,
            "reversible": bool(i.reversible), "calls": calls,
            "mpixel_s": round(px / float(np.mean(lat)) / 1e6, 2), "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
            "resident_ms": round(res_s * 1e3, 3), "resident_mpixel_s": round(px / res_s / 1e6, 1),
            "bit_exact": src is not None}, out


def bench_reduced(tca, cs, steps):
    """The headline codestream decoded at reduced resolution (the k finest DWT
    levels discarded): resident decode time, output Mpixel/s, and the speed-up
    over the full-resolution decode.  Outputs checked against the oracle's
    reduced decode first (within 1 LSB: the stream is lossy)."""
    from oracle import oracle
    res = {}
    full_ms = None
    for k in (0, 1, 2):
        dec = tca.Decoder()
        dec.set_reduce(k)
        out = dec.decode(cs)
        if k:
            ref = oracle.decode(cs, reduce=k).transpose(1, 2, 0)
            d = int(np.abs(out.astype(np.int32) - ref).max())
            if d > 1:
                raise SystemExit(f"reduced decode (k={k}) differs from the oracle by {d}")
        dec.prepare([cs])
        dec.run()
        dec.sync()
        ms = timed_runs(dec, steps) * 1e3
        dec.close()
        px = out.shape[0] * out.shape[1]
        if k == 0:
            full_ms = ms
            continue
        res[f"reduce{k}"] = {"width": out.shape[1], "height": out.shape[0], "ms": round(ms, 3),
                             "output_mpixel_s": round(px / ms / 1e3, 1), "speedup_vs_full": round(full_ms / ms, 2),
                             "max_diff_vs_oracle": d}
    res["note"] = ("cfg2 codestream, resident in HBM; tier-1 skips the discarded levels' code-blocks, but its "
                   "time stays bound by the lowest-resolution block, which a reduced decode still needs")
    return res


def bench_forward(tca, steps):
    """GPU forward transform (encoder front end: DC shift + forward ICT/RCT fused
    into the first 9/7 or 5/3 level, then the other levels) of the cfg2 and cfg3
    source images, resident in HBM.  Checked against the oracle's forward DWT
    first (9/7 within FP32 rounding, 5/3 exact).  Roofline of the fused first
    level: bytes = the image read once + 4 B per sample and component written."""
    from oracle import oracle
    from thousandthchicken_amd import synth
    res = {}
    for name, (h, w, c, bits, levels, irr, seed) in {
            "cfg2_2048_rgb8_97": (2048, 2048, 3, 8, 5, True, 2),
            "cfg3_3840x2160_rgb12_53": (2160, 3840, 3, 12, 6, False, 3)}.items():
        img = synth.image(h, w, c, seed, bits)
        f = tca.Forward()
        f.prepare(w, h, c, bits, levels, irr, True)
        f.upload(img)
        f.run()
        f.sync()
        got = f.download()
        # oracle check on the luma plane (the other planes share every kernel path)
        x = img.astype(np.float64 if irr else np.int64) - (1 << (bits - 1))
        if irr:
            y = 0.299 * x[..., 0] + 0.587 * x[..., 1] + 0.114 * x[..., 2]
            ref = oracle.fdwt97(y, 0, 0, levels)
            d = float(np.abs(got[0] - ref).max())
            if d > 2e-4 * max(1.0, float(np.abs(ref).max())):
                raise SystemExit(f"forward 9/7 differs from the oracle by {d}")
        else:
            y = (x[..., 0] + 2 * x[..., 1] + x[..., 2]) >> 2
            d = float(np.abs(got[0] - oracle.fdwt53(y.astype(np.int32), 0, 0, levels)).max())
            if d:
                raise SystemExit(f"forward 5/3 differs from the oracle by {d}")
        for _ in range(3):
            f.run_timed()
        firsts, totals = [], []
        for _ in range(steps):
            a, b = f.run_timed()
            firsts.append(a)
            totals.append(b)
        f.close()
        first_ms, total_ms = float(np.median(firsts)), float(np.median(totals))
        in_b = 2 if bits > 8 else 1
        first_bytes = w * h * c * (in_b + 4)
        gbs = first_bytes / (first_ms * 1e-3) / 1e9
        res[name] = {"width": w, "height": h, "components": c, "bits": bits, "levels": levels,
                     "wavelet": "9/7+ICT" if irr else "5/3+RCT", "ms": round(total_ms, 4),
                     "mpixel_s": round(w * h / total_ms / 1e3, 1), "first_level_ms": round(first_ms, 4),
                     "max_diff_vs_oracle_luma": d,
                     "roofline_first_level": {"bound": "hbm", "kernel": "fdwt_first_kernel", "achieved": round(gbs, 1),
                                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                              "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes": first_bytes}}
    return res


def bench_region(tca, steps):
    """Region decode of the cfg3 codestream (3840x2160 12-bit lossless, twelve
    1024x1024 tiles): resident decode time of windows against the full image,
    with the code-blocks decoded.  Each window is checked against the same
    window of the full decode first (bit-exact)."""
    from thousandthchicken_amd import workloads
    cs = workloads.cfg3()
    full_dec = tca.Decoder()
    full = full_dec.decode(cs)
    full_dec.prepare([cs])
    full_dec.run()
    full_dec.sync()
    full_ms = timed_runs(full_dec, steps) * 1e3
    full_blocks = full_dec.stats().cblks
    full_dec.close()
    res = {"full": {"ms": round(full_ms, 3), "cblks": full_blocks}}
    for name, (x, y, w, h) in {"center_512": (1664, 824, 512, 512), "quarter_1920x1080": (0, 0, 1920, 1080),
                               "strip_3840x64": (0, 1000, 3840, 64)}.items():
        dec = tca.Decoder()
        dec.set_region(x, y, w, h)
        out = dec.decode(cs)
        if not np.array_equal(out, full[y:y + h, x:x + w]):
            raise SystemExit(f"region {name} differs from the full decode")
        dec.prepare([cs])
        dec.run()
        dec.sync()
        ms = timed_runs(dec, steps) * 1e3
        res[name] = {"window": [x, y, w, h], "ms": round(ms, 3), "cblks": dec.stats().cblks,
                     "speedup_vs_full": round(full_ms / ms, 2), "output_mpixel_s": round(w * h / ms / 1e3, 1)}
        dec.close()
    res["note"] = ("cfg3 codestream resident in HBM; only the code-blocks whose lifting support reaches the window "
                   "are decoded, the inverse DWT still runs over every tile")
    return res


def profile_mode(tca, args):
    """A few steps of the workload (under the packing given on the command line,
    no autotuning, so that the trace holds only the decode's own dispatches)
    plus one device copy of known size, for scaling the traffic counters."""
    from thousandthchicken_amd import workloads
    make = {"cfg1_lossless_512": lambda: [workloads.cfg1()], "cfg2_lossy_2048": lambda: [workloads.cfg2()],
            "cfg3_lossless_4k_tiled": lambda: [workloads.cfg3()], "cfg4_mixed_256": lambda: workloads.cfg4(threads=16),
            "cfg5_large_16bit": lambda: [workloads.cfg5_large(5)]}
    streams = make[args.workload]()
    decs = []
    for _ in range(max(1, args.profile_in_flight)):
        d = tuned_decoder(tca, args)
        d.prepare(streams)
        decs.append(d)
    dec = decs[0]
    for k in range(args.steps * len(decs)):
        decs[k % len(decs)].run()
    for d in decs:
        d.sync()
    s = dec.stats()
    n = 256 << 20
    L = tca.lib()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    tca.check(L.tc_buf_alloc(dec.device.handle, n, ctypes.byref(a)))
    tca.check(L.tc_buf_alloc(dec.device.handle, n, ctypes.byref(b)))
    st = L.tc_decoder_stream(dec.handle)
    # the copied bytes' values do not matter to the counters: no fill kernel
    # runs here, so the decode's own fill dispatches stay alone in the trace
    tca.check(L.tc_buf_copy_d2d(st, b, 0, a, 0, n))
    tca.check(L.tc_stream_sync(st))
    L.tc_buf_free(a)
    L.tc_buf_free(b)
    print(json.dumps({"profile": args.workload, "steps": args.steps, "copy_bytes": n,
                      "blocks_per_wave": {"small": s.bpw_small, "large": s.bpw_large},
                      "t1_launches": [{"blocks_per_wave": s.t1_launch_bpw[g], "blocks": s.t1_launch_blocks[g],
                                       "bytes": s.t1_launch_bytes[g]} for g in range(s.t1_launches)]}), flush=True)


def bench_cfg3(tca, steps, args):
    """The 3840x2160 12-bit image (twelve 1024x1024 tiles, 32x32 blocks):
    batched from a graph against the per-image entry; bit-exact first."""
    from thousandthchicken_amd import workloads
    cs, src = workloads.cfg3(), workloads.source_image("cfg3")
    dec = tuned_decoder(tca, args)
    dec.prepare([cs])
    tuning = None if args.no_autotune else dec.autotune(reps=2).as_dict()
    dec.run()
    dec.sync()
    check_equal_source(dec.fetch(0), src, "cfg3")
    dec.capture_graph()
    dec.launch_graph()
    graph = timed_runs(dec, steps, graph=True)
    s = dec.stats()
    dec.close()
    one = tca.Decoder()
    one.decode(cs)
    t0 = time.perf_counter()
    for _ in range(3):
        one.decode(cs)
    per = (time.perf_counter() - t0) / 3
    one.close()
    return {"mpixel_s_graph": round(3840 * 2160 / graph / 1e6, 1), "ms": round(graph * 1e3, 3), "cblks": s.cblks,
            "per_image_mpixel_s": round(3840 * 2160 / per / 1e6, 1), "bit_exact": True, "autotuned": tuning}


def bench_served(tca, small, large, loads, args, serial_requests=300):
    """Poisson loads through the server.  Each load offers min(5 s x rate,
    3000) requests; a load counts as sustained when the achieved rate is
    within 5% of the offered rate or the latency does not grow over the run
    (last fifth of the requests against the second fifth)."""
    from thousandthchicken_amd import serve
    n_max = max(min(int(5 * r), 3000) for r in loads)
    order = serve.mix(small, large, n_max, 0.1, seed=11)
    seq = [small[i] if i >= 0 else large[-1 - i] for i in order]
    keep = set(range(0, n_max, max(1, n_max // 24)))
    dec = tca.Decoder()
    for cs in seq[:3]:
        dec.decode(cs)
    ser = serve.serial(dec, seq[:serial_requests])
    ref = {i: dec.decode(seq[i]) for i in keep}
    dec.close()
    out = {"large_fraction": 0.1, "serial_per_image": ser, "loads": []}
    srv = tca.Server(slots=args.slots, max_batch=args.max_batch, bpw_small=args.bpw_small, bpw_large=args.bpw_large,
                     max_batch_bytes=args.batch_mib << 20, overlap_kernels=args.overlap)
    # warm: the slots' resident device buffers and pinned buffers reach their working size
    serve.run(srv, seq[:1000], max(loads), seed=1)
    best = None
    for rate in loads:
        n = min(int(5 * rate), 3000)
        srv.reset_stats()
        r = serve.run(srv, seq[:n], rate, seed=int(rate), keep={i for i in keep if i < n})
        for i, img in r.pop("images").items():
            d = int(np.abs(img.astype(np.int32) - ref[i].reshape(img.shape).astype(np.int32)).max())
            if d > (0 if tca.probe(seq[i]).reversible else 1):
                raise SystemExit(f"served request {i} differs from its per-image decode by {d}")
        r["speedup_vs_serial"] = round(r["mpixel_s"] / ser["mpixel_s"], 2)
        # keeping up: the whole run's rate is within 5 % of the offer, or
        # within 20 % of it with the latency of the last fifth of the
        # requests not grown past 1.5x that of the second fifth (a short run's
        # drain tail, the last batches' latency after the last arrival, costs
        # up to ~15 % of the achieved rate at 2000 req/s)
        r["sustained"] = bool(r["achieved_rps"] >= 0.95 * rate or
                              (r["achieved_rps"] >= 0.8 * rate and
                               r["p50_ms_last_fifth"] <= 1.5 * r["p50_ms_second_fifth"]))
        out["loads"].append(r)
        log(f"served {rate} req/s: {r}")
        if r["sustained"]:
            best = r
    srv.close()
    out["max_sustained"] = None if best is None else {
        k: best[k] for k in ("offered_rps", "achieved_rps", "mpixel_s", "p50_ms", "p99_ms", "speedup_vs_serial")}
    out["peak"] = max(out["loads"], key=lambda r: r["mpixel_s"])["mpixel_s"]
    out["warm"] = "one untimed pass of 1000 requests before the timed loads"
    out["outputs"] = (f"{len(keep)} sampled requests checked against per-image decode "
                      "(lossless identical, lossy within 1 LSB)")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="cfg4_mixed_256",
                    choices=["cfg1_lossless_512", "cfg2_lossy_2048", "cfg3_lossless_4k_tiled", "cfg4_mixed_256",
                             "cfg5_large_16bit"],
                    help="workload of --profile")
    ap.add_argument("--quick", action="store_true", help="headline only")
    ap.add_argument("--no-served", action="store_true", help="skip the serving runs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile", action="store_true", help="workload steps and a calibration copy only")
    ap.add_argument("--forward-only", action="store_true",
                    help="only the GPU forward transform (encoder front end) leg")
    ap.add_argument("--region-only", action="store_true", help="only the region-decode leg")
    ap.add_argument("--served-only", action="store_true", help="only the serving leg")
    ap.add_argument("--loads", default="100,250,500,1000,2000,3000,5000", help="offered loads (req/s) of the serving run")
    ap.add_argument("--slots", type=int, default=2, help="server slots")
    ap.add_argument("--max-batch", type=int, default=256, help="server: most requests per batch")
    ap.add_argument("--batch-mib", type=int, default=1024, help="server: most codestream MiB per batch")
    ap.add_argument("--overlap", action="store_true", help="server: kernels of different slots may run together")
    ap.add_argument("--bpw-large", type=int, default=0, help="blocks per wave, blocks up to 64x64 (0 auto, -64/-32/-16 lockstep)")
    ap.add_argument("--bpw-small", type=int, default=0, help="blocks per wave, blocks up to 32x32 (0 auto, -64/-32/-16 lockstep)")
    ap.add_argument("--t1-kernel", type=int, default=0, help="one-wave-per-block kernel: 0 stripe-column, 1 row-mask")
    ap.add_argument("--dwt-shape", type=int, default=-1, help="inverse-DWT tile shape (tc_idwt_tile_shape; -1 default)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default packing and DWT tile shape")
    ap.add_argument("--in-flight", default="1,2,4,8",
                    help="batches in flight tried for the headline (each its own pipeline and stream)")
    ap.add_argument("--profile-in-flight", type=int, default=1,
                    help="--profile: batches in flight (each its own pipeline, launched round robin)")
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="GPU_MAX_HW_QUEUES for this process (set before HIP starts; 0 leaves it)")
    args = ap.parse_args()
    if args.hw_queues > 0:
        # several pipelines in flight use several streams each; with HIP's
        # default of 4 hardware queues, streams share queues and their work
        # serialises behind each other
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(args.hw_queues, 32))
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import thousandthchicken_amd as tca
    from thousandthchicken_amd import workloads
    if args.profile:
        return profile_mode(tca, args)
    if args.forward_only:
        print(json.dumps({"forward": bench_forward(tca, args.steps)}), flush=True)
        return
    if args.region_only:
        print(json.dumps({"region": bench_region(tca, args.steps)}), flush=True)
        return
    t = time.perf_counter()
    streams = workloads.cfg4(threads=16)
    log(f"256-image batch generated in {time.perf_counter() - t:.1f} s")
    if args.served_only:
        large = [workloads.cfg5_large(5), workloads.cfg5_large(6)]
        print(json.dumps({"served": bench_served(tca, streams, large,
                                                 [float(x) for x in args.loads.split(",") if x], args)}), flush=True)
        return
    lay = workloads.cfg4_layout()

    # ---- headline: the 256-image batch, one launch per stage, from a graph ----
    serial_dec = tca.Decoder()
    for cs in streams[:4]:
        serial_dec.decode(cs)
    outs, px = [], 0
    t0 = time.perf_counter()
    for cs in streams:
        o = serial_dec.decode(cs)
        outs.append(o)
        px += o.shape[0] * o.shape[1]
    serial_s = time.perf_counter() - t0
    serial_dec.close()
    dec = tuned_decoder(tca, args)
    dec.prepare(streams)
    t0 = time.perf_counter()
    for _ in range(3):
        dec.prepare(streams)
    prep_ms = (time.perf_counter() - t0) / 3 * 1e3
    tuning = None if args.no_autotune else dec.autotune(reps=1).as_dict()
    dec.run()
    dec.sync()
    batch_outs = [dec.fetch(i) for i in range(len(streams))]
    for i in range(len(streams)):
        d = int(np.abs(batch_outs[i].astype(np.int32) - outs[i].astype(np.int32)).max())
        if d > (0 if lay[i]["lossless"] else 1):
            raise SystemExit(f"batch image {i}: batched differs from per-image by {d}")
    dec.capture_graph()
    for _ in range(args.warmup):
        dec.launch_graph()
    graph = timed_runs(dec, args.steps, graph=True)
    eager = timed_runs(dec, max(3, args.steps // 4))
    st, launch_ms = stage_times(dec)
    s = dec.stats()
    dec.close()
    lossless = [lay[i]["lossless"] for i in range(len(streams))]
    flight, flight_tried, flight_s = in_flight_leg(tca, streams, outs, lossless, px, args, tuning)
    if flight_s > graph:  # one in flight (the latency-tuned graph) stays the faster form
        flight = {"bpw_small": s.bpw_small, "bpw_large": s.bpw_large, "in_flight": 1,
                  "ms_per_batch": round(graph * 1e3, 3), "mpixel_s": round(px / graph / 1e6, 1)}
        flight_s = graph
    launches = t1_launches(s, launch_ms, args)
    # the dominant kernel: the longest tier-1 launch
    dom = max(launches, key=lambda x: x["ms"])
    kname = dom["kernel"]
    t1_bytes = dom["bytes"]
    t1_gbs = t1_bytes / (dom["ms"] * 1e-3) / 1e9
    dwt_gbs = (s.dwt_bytes + s.final_bytes) / ((st["dwt_ms"] + st["final_ms"]) * 1e-3) / 1e9
    tr = traffic_record(kname, "cfg4_mixed_256")
    line = {
        "metric": "decoded Mpixel/s on 1 MI355X",
        "value": round(px / flight_s / 1e6, 1),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(flight_s * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+fp32",
        "data": "synthetic (seeded recipe, coded by the package encoder)",
        "config": {"workload": "cfg4_mixed_256", "images": len(streams), "mpixel": round(px / 1e6, 2),
                   "codestream_bytes": sum(map(len, streams)), "cblks": s.cblks,
                   "mix": "sides log-uniform 256-2048, 1 or 3 components, 8-bit, half lossless 5/3, "
                          "half lossy 9/7 0.5-2 bpp, 32x32 or 64x64 blocks",
                   "form": "batched, one launch per stage per batch, graph replay, "
                           f"{flight['in_flight']} batch(es) in flight on their own pipelines and streams",
                   "in_flight": flight["in_flight"], "inputs": "codestreams resident in HBM"},
        "in_flight": {"chosen": flight, "tried": flight_tried,
                      "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
                      "note": "a step is one whole decode of the batch; copies of the prepared batch on their own "
                              "pipelines are launched round robin, so one batch's tier-1 tail overlaps the next "
                              "batch's work; every copy's outputs are checked against per-image decode. "
                              "stages_ms, tier1_launches and roofline below are one batch alone"},
        "stages_ms": {"fill": round(st["fill_ms"], 3), "tier1": round(st["t1_ms"], 3),
                      "idwt_levels": round(st["dwt_ms"], 3), "final_level": round(st["final_ms"], 3),
                      "total_eager": round(st["total_ms"], 3)},
        "tier1_launches": launches,
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(t1_gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(t1_gbs / HBM_PEAK_GBS, 5),
                     "traffic": tr["hbm_bytes_per_launch"] if tr else None, "algorithmic_bytes": t1_bytes,
                     "launch_ms": dom["ms"], "blocks": dom["blocks"],
                     "note": "tier-1 is bound by the serial MQ symbol chains of its blocks, not by HBM; bytes = "
                             "codeword bytes + 4 B per coefficient of the launch's blocks, over the launch's own "
                             "time (HIP events on the stream it runs on)"},
        "roofline_valu": roofline_valu(flight, "cfg4_mixed_256"),
        "roofline_dwt": {"bound": "hbm", "achieved": round(dwt_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(dwt_gbs / HBM_PEAK_GBS, 5)},
        "blocks_per_wave": {"small": s.bpw_small, "large": s.bpw_large},
        "heavy_threshold": {"small": s.heavy_small, "large": s.heavy_large},
        "autotuned": tuning,
        "batched": {"cfg4_mixed_256": {
            "mpixel_s_in_flight": round(px / flight_s / 1e6, 1),
            "mpixel_s_graph": round(px / graph / 1e6, 1), "mpixel_s_eager": round(px / eager / 1e6, 1),
            "ms_per_batch_one_in_flight": round(graph * 1e3, 3), "prepare_ms": round(prep_ms, 1),
            "per_image_serial_mpixel_s": round(px / serial_s / 1e6, 1),
            "speedup_vs_per_image": round(serial_s / flight_s, 2),
            "speedup_one_in_flight_vs_per_image": round(serial_s / graph, 2),
            "outputs": "lossless identical to per-image decode, lossy within 1 LSB"}},
    }
    if not args.quick:
        lossy = [i for i in range(len(streams)) if not lay[i]["lossless"]]
        with cf.ThreadPoolExecutor(cpu_threads()) as ex:
            parts = list(ex.map(lambda i: lossy_errors(batch_outs[i], streams[i]), lossy))
        acc = {"cfg4_lossy_images": dict(error_summary(parts), images=len(lossy))}
        del batch_outs, outs
        per = {}
        per["cfg1_lossless_512"], _ = per_image_leg(tca, workloads.cfg1(), workloads.source_image("cfg1"), 50, "cfg1")
        cs2 = workloads.cfg2()
        per["cfg2_lossy_2048"], out2 = per_image_leg(tca, cs2, None, 30, "cfg2")
        acc["cfg2_lossy_2048"] = error_summary([lossy_errors(out2, cs2)])
        if acc["cfg2_lossy_2048"]["max_lsb_vs_rounded_fp64"] > 1:
            raise SystemExit("cfg2 differs from the FP64 oracle by more than 1 LSB")
        large = [workloads.cfg5_large(5), workloads.cfg5_large(6)]
        per["cfg5_large_16bit"], _ = per_image_leg(tca, large[0], workloads.source_image("cfg5_large"), 10,
                                                   "cfg5 large", resident_steps=5)
        line["per_image"] = per
        line["accuracy"] = acc
        line["batched"]["cfg3_lossless_4k_tiled"] = bench_cfg3(tca, max(3, args.steps // 4), args)
        line["reduced"] = bench_reduced(tca, cs2, max(5, args.steps // 2))
        line["forward"] = bench_forward(tca, max(5, args.steps // 2))
        line["region"] = bench_region(tca, max(5, args.steps // 4))
        if not args.no_served:
            line["served"] = bench_served(tca, streams, large, [float(x) for x in args.loads.split(",") if x], args)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(streams)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
