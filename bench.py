"""Decode throughput benchmark (decoded Mpixel/s on one MI355X).

Headline workload: one 2048x2048 3-component 8-bit image from the seeded
recipe (seed 2), coded lossy with ICT + 5-level 9/7, expounded quantization,
64x64 code-blocks, one tile, one layer truncated to about 1.0 bpp (the
package's encoder).  A step decodes it from the device-resident codestream to
the interleaved 8-bit image in HBM: fill, tier-1, inverse DWT levels, fused
final level (dequantization, ICT, DC shift, rounding, packing).

Beside the headline line the same JSON object carries (unless --quick):
  per_image  the same image through the per-image entry (host codestream in,
             host image out, blocking), one call after the other;
  batched    the mixed batch of 256 images (and the 3840x2160 12-bit tiled
             image) decoded with one launch per stage, eager and from a
             captured graph, against the same images decoded one by one;
  served     Poisson arrivals (90% mixed-batch images, 10% 7680x4320 16-bit)
             at several offered loads through the decode server, with p50/p99
             latency, against serial per-image decode of the same requests.

Prints one JSON line.
"""
import argparse
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


def workload(name):
    from thousandthchicken_amd import workloads
    if name == "cfg1_lossless_512":
        return workloads.cfg1(), workloads.source_image("cfg1")
    if name == "cfg2_lossy_2048":
        return workloads.cfg2(), None
    if name == "cfg3_lossless_4k_tiled":
        return workloads.cfg3(), workloads.source_image("cfg3")
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(cs, pixels, budget_s=20.0):
    """The oracle (scalar C, one core) decoding the same codestream."""
    from oracle import oracle
    t0 = time.perf_counter()
    n = 0
    while True:
        oracle.decode(cs)
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s / 4 or n >= 3:
            break
    return {"value": round(pixels * n / el / 1e6, 3), "unit": "Mpixel/s", "cores": 1, "kind": "port",
            "sample": f"{n} full decode(s) of the benchmark codestream by the C oracle, {el:.1f} s"}


def check_output(out, cs, src, what):
    if src is not None:
        if not np.array_equal(out.reshape(src.shape), src):
            raise SystemExit(f"{what}: decode is not bit-exact")
        return
    from oracle import oracle
    ref = oracle.decode(cs).transpose(1, 2, 0)
    dmax = int(np.abs(out.astype(np.int32) - ref).max())
    if dmax > 1:
        raise SystemExit(f"{what}: decode differs from the oracle by {dmax} LSB")


def stage_times(dec, reps=10):
    acc = np.zeros(5)
    for _ in range(reps):
        dec.run(timed=True)
        dec.sync()
        s = dec.stats()
        acc += [s.fill_ms, s.t1_ms, s.dwt_ms, s.final_ms, s.total_ms]
    return acc / reps


def timed_runs(dec, steps, graph=False):
    dec.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        dec.launch_graph() if graph else dec.run()
    dec.sync()
    return (time.perf_counter() - t0) / steps


def t1_kernel_name(stats, args):
    bpw = stats.bpw_large or stats.bpw_small
    if bpw < 0:
        return "t1_lockstep_kernel"
    if bpw > 1:
        return f"t1_decode_kernel<{bpw}>"
    return "t1_stripe_kernel" if args.t1_kernel == 0 else "t1_wave_kernel"


def traffic_record(kernel, workload):
    """HBM bytes per tier-1 launch measured with FETCH_SIZE / WRITE_SIZE
    (profiles/traffic_<kernel>_<workload>.json)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel.split('<')[0]}_{workload}.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def tuned_decoder(tca, args):
    dec = tca.Decoder()
    dec.set_blocks_per_wave(False, args.bpw_large)
    dec.set_blocks_per_wave(True, args.bpw_small)
    dec.set_t1_kernel(args.t1_kernel)
    return dec


def bench_batched(tca, steps, args):
    from thousandthchicken_amd import workloads
    res = {}
    t = time.perf_counter()
    streams = workloads.cfg4(threads=16)
    log(f"mixed batch generated in {time.perf_counter() - t:.1f} s")
    lay = workloads.cfg4_layout()
    px = 0
    serial_dec = tca.Decoder()
    outs = []
    for cs in streams[:4]:
        serial_dec.decode(cs)
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
    tuning4 = None if args.no_autotune else dec.autotune(reps=1).as_dict()
    dec.run()
    dec.sync()
    lossy_diff = 0
    for i in range(len(streams)):
        d = int(np.abs(dec.fetch(i).astype(np.int32) - outs[i]).max())
        if d > (0 if lay[i]["lossless"] else 1):
            raise SystemExit(f"mixed batch image {i}: batched differs from per-image by {d}")
        lossy_diff = max(lossy_diff, d)
    st = stage_times(dec, 3)
    eager = timed_runs(dec, steps)
    dec.capture_graph()
    dec.launch_graph()
    graph = timed_runs(dec, steps, graph=True)
    s = dec.stats()
    res["cfg4_mixed_256"] = {
        "images": len(streams), "mpixel": round(px / 1e6, 2), "cblks": s.cblks,
        "mpixel_s_eager": round(px / eager / 1e6, 1), "mpixel_s_graph": round(px / graph / 1e6, 1),
        "ms_per_batch": round(graph * 1e3, 3),
        "prepare_ms": round(prep_ms, 1),
        "stages_ms": {"fill": round(st[0], 3), "tier1": round(st[1], 3), "idwt_levels": round(st[2], 3),
                      "final_level": round(st[3], 3)},
        "per_image_serial_mpixel_s": round(px / serial_s / 1e6, 1),
        "speedup_vs_per_image": round(serial_s / graph, 2),
        "blocks_per_wave": {"small": s.bpw_small, "large": s.bpw_large},
        "autotuned": tuning4,
        "outputs": "lossless identical to per-image decode",
        "lossy_max_diff_vs_per_image": lossy_diff}
    dec.close()
    # one large tiled image (12-bit, 1024x1024 tiles, 32x32 blocks)
    cs, src = workload("cfg3_lossless_4k_tiled")
    dec = tuned_decoder(tca, args)
    dec.prepare([cs])
    tuning3 = None if args.no_autotune else dec.autotune(reps=2).as_dict()
    dec.run()
    dec.sync()
    check_output(dec.fetch(0), cs, src, "cfg3")
    dec.capture_graph()
    dec.launch_graph()
    graph = timed_runs(dec, steps, graph=True)
    one = tca.Decoder()
    one.decode(cs)
    t0 = time.perf_counter()
    for _ in range(3):
        one.decode(cs)
    per = (time.perf_counter() - t0) / 3
    s = dec.stats()
    res["cfg3_lossless_4k_tiled"] = {
        "mpixel_s_graph": round(3840 * 2160 / graph / 1e6, 1), "ms": round(graph * 1e3, 3), "cblks": s.cblks,
        "per_image_mpixel_s": round(3840 * 2160 / per / 1e6, 1), "bit_exact": True,
        "autotuned": tuning3}
    dec.close()
    one.close()
    return res


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


def bench_served(tca, loads, n_req, args):
    from thousandthchicken_amd import serve, workloads
    t = time.perf_counter()
    small = workloads.cfg4(threads=16)
    large = [workloads.cfg5_large(5), workloads.cfg5_large(6)]
    log(f"serving pool generated in {time.perf_counter() - t:.1f} s")
    order = serve.mix(small, large, n_req, 0.1, seed=11)
    seq = [small[i] if i >= 0 else large[-1 - i] for i in order]
    keep = set(range(0, n_req, max(1, n_req // 16)))
    dec = tca.Decoder()
    for cs in seq[:3]:
        dec.decode(cs)
    ser = serve.serial(dec, seq)
    ref = {i: dec.decode(seq[i]) for i in keep}
    dec.close()
    out = {"requests": n_req, "large_fraction": 0.1, "serial_per_image": ser, "loads": []}
    # one server for every load, warmed with one untimed pass so that its
    # resident buffers have reached their working size
    srv = tca.Server(slots=3, max_batch=256, bpw_small=args.bpw_small, bpw_large=args.bpw_large)
    serve.run(srv, seq, max(loads), seed=1)
    for rate in loads:
        srv.reset_stats()
        r = serve.run(srv, seq, rate, seed=int(rate), keep=keep)
        for i, img in r.pop("images").items():
            d = int(np.abs(img.astype(np.int32) - ref[i]).max())
            if d > (0 if tca.probe(seq[i]).reversible else 1):
                raise SystemExit(f"served request {i} differs from its per-image decode by {d}")
        r["speedup_vs_serial"] = round(r["mpixel_s"] / ser["mpixel_s"], 2)
        out["loads"].append(r)
        log(f"served {rate} req/s: {r}")
    srv.close()
    out["warm"] = "one untimed pass of the same requests before the timed loads"
    out["outputs"] = (f"{len(keep)} sampled requests checked against per-image decode "
                      "(lossless identical, lossy within 1 LSB)")
    return out


def profile_mode(tca, args):
    """A few steps of the workload plus a device copy of known size, for counters."""
    if args.workload == "cfg4_mixed_256":
        from thousandthchicken_amd import workloads
        streams = workloads.cfg4(threads=16)
    else:
        streams = [workload(args.workload)[0]]
    dec = tuned_decoder(tca, args)
    dec.prepare(streams)
    for _ in range(args.steps):
        dec.run()
    dec.sync()
    n = 256 << 20
    L = tca.lib()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    tca.check(L.tc_buf_alloc(dec.device.handle, n, ctypes.byref(a)))
    tca.check(L.tc_buf_alloc(dec.device.handle, n, ctypes.byref(b)))
    st = L.tc_decoder_stream(dec.handle)
    tca.check(L.tc_buf_fill32(st, a, 0, 1, n // 4))
    for _ in range(3):
        tca.check(L.tc_buf_copy_d2d(st, b, 0, a, 0, n))
    tca.check(L.tc_stream_sync(st))
    L.tc_buf_free(a)
    L.tc_buf_free(b)
    print(json.dumps({"profile": args.workload, "steps": args.steps, "copy_bytes": n}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="cfg2_lossy_2048",
                    choices=["cfg1_lossless_512", "cfg2_lossy_2048", "cfg3_lossless_4k_tiled", "cfg4_mixed_256"],
                    help="headline workload (cfg4_mixed_256: --profile only)")
    ap.add_argument("--quick", action="store_true", help="headline workload only (no batched/served)")
    ap.add_argument("--no-served", action="store_true", help="skip the serving runs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile", action="store_true", help="workload steps and a calibration copy only")
    ap.add_argument("--forward-only", action="store_true",
                    help="only the GPU forward transform (encoder front end) leg")
    ap.add_argument("--region-only", action="store_true", help="only the region-decode leg")
    ap.add_argument("--loads", default="50,100,1000,5000", help="offered loads (req/s) of the serving run")
    ap.add_argument("--requests", type=int, default=300, help="requests per serving run")
    ap.add_argument("--bpw-large", type=int, default=0, help="blocks per wave, blocks up to 64x64 (0 auto, -64/-32 lockstep)")
    ap.add_argument("--bpw-small", type=int, default=0, help="blocks per wave, blocks up to 32x32 (0 auto, -64/-32 lockstep)")
    ap.add_argument("--t1-kernel", type=int, default=0, help="one-wave-per-block kernel: 0 stripe-column, 1 row-mask")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default packing and DWT tile shape")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import thousandthchicken_amd as tca
    if args.profile:
        return profile_mode(tca, args)
    if args.forward_only:
        print(json.dumps({"forward": bench_forward(tca, args.steps)}), flush=True)
        return
    if args.region_only:
        print(json.dumps({"region": bench_region(tca, args.steps)}), flush=True)
        return

    cs, src = workload(args.workload)
    info = tca.probe(cs)
    pixels = info.width * info.height
    dec = tuned_decoder(tca, args)
    dec.prepare([cs])
    tuning = None if args.no_autotune else dec.autotune(reps=2).as_dict()
    for _ in range(args.warmup):
        dec.run()
    dec.sync()
    check_output(dec.fetch(0), cs, src, args.workload)

    step = timed_runs(dec, args.steps)
    ms = step * 1e3
    st = stage_times(dec)
    s = dec.stats()
    # tier-1 algorithmic bytes per launch: codeword bytes read + 4 B per coefficient written
    # (the coefficients of the blocks that have coding passes; the rest stay as filled)
    t1_bytes = s.codeword_bytes + 4 * s.t1_samples
    t1_gbs = t1_bytes / (st[1] * 1e-3) / 1e9
    dwt_gbs = (s.dwt_bytes + s.final_bytes) / ((st[2] + st[3]) * 1e-3) / 1e9 if st[2] + st[3] > 0 else 0.0
    tr = traffic_record(t1_kernel_name(s, args), args.workload)
    traffic = tr["hbm_bytes_per_launch"] if tr else None

    # plain per-image entry (host in, host out), one call after the other
    per = tca.Decoder()
    per.decode(cs)
    t0 = time.perf_counter()
    n_per = max(5, args.steps // 5)
    for _ in range(n_per):
        per.decode(cs)
    per_s = (time.perf_counter() - t0) / n_per
    per.close()
    # the same with zero-copy output: the final kernel writes pinned host memory
    # and the caller reads it in place (codestream still uploaded per call)
    zc = tca.Decoder()
    zc.set_host_output(True)
    zc.prepare([cs])
    zc.run()
    zc.sync()
    if not np.array_equal(zc.output_view(0), dec.fetch(0)):
        raise SystemExit("zero-copy host output differs from the device output")
    t0 = time.perf_counter()
    for _ in range(n_per):
        zc.prepare([cs])
        zc.run()
        zc.sync()
        zc.output_view(0)
    zc_s = (time.perf_counter() - t0) / n_per
    zc.close()

    line = {
        "metric": "decoded Mpixel/s on 1 MI355X",
        "value": round(pixels / step / 1e6, 2),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+fp32",
        "data": "synthetic (seeded recipe, coded by the package encoder)",
        "config": {"workload": args.workload, "width": info.width, "height": info.height,
                   "components": info.ncomp, "reversible": bool(info.reversible), "levels": info.nlevels,
                   "codeblock": info.cbw, "codestream_bytes": len(cs),
                   "bpp": round(len(cs) * 8 / pixels, 3), "inputs": "codestream resident in HBM"},
        "stages_ms": {"fill": round(st[0], 4), "tier1": round(st[1], 4), "idwt_levels": round(st[2], 4),
                      "final_level": round(st[3], 4), "total": round(st[4], 4)},
        "roofline": {"bound": "hbm", "kernel": t1_kernel_name(s, args), "achieved": round(t1_gbs, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(t1_gbs / HBM_PEAK_GBS, 5),
                     "traffic": traffic, "algorithmic_bytes": t1_bytes,
                     "note": "tier-1 is bound by the serial MQ symbol chain of the heaviest code-block, "
                             "not by HBM; bytes = codewords + 4 B per coefficient"},
        "roofline_dwt": {"bound": "hbm", "achieved": round(dwt_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(dwt_gbs / HBM_PEAK_GBS, 5)},
        "blocks_per_wave": {"small": s.bpw_small, "large": s.bpw_large},
        "autotuned": tuning,
        "cblks": s.cblks,
        "per_image": {"mpixel_s": round(pixels / per_s / 1e6, 2), "ms": round(per_s * 1e3, 3),
                      "note": "host codestream in, host image out, PCIe included"},
        "per_image_host_output": {"mpixel_s": round(pixels / zc_s / 1e6, 2), "ms": round(zc_s * 1e3, 3),
                                  "note": "same, the final kernel writing pinned host memory (no output copy)"},
    }
    dec.close()
    if not args.quick:
        line["reduced"] = bench_reduced(tca, cs, max(5, args.steps // 2))
        line["forward"] = bench_forward(tca, max(5, args.steps // 2))
        line["region"] = bench_region(tca, max(5, args.steps // 5))
        line["batched"] = bench_batched(tca, max(3, args.steps // 5), args)
        loads = [float(x) for x in args.loads.split(",") if x]
        if not args.no_served:
            line["served"] = bench_served(tca, loads, args.requests, args)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(cs, pixels)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
