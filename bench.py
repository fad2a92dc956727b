"""Decode throughput benchmark (decoded Mpixel/s on one MI355X).

Default workload: one 2048x2048 3-component 8-bit image from the seeded
recipe (seed 2), coded lossy with ICT + 5-level 9/7, expounded quantization,
64x64 code-blocks, one tile, one layer truncated to about 1.0 bpp.  A step
decodes it from the device-resident codestream to the interleaved 8-bit
image in HBM (fill, tier-1, inverse DWT levels, fused final level).

Prints one JSON line (rank 0).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (height, width, ncomp, bits, seed, lossless, levels, cb, bpp)
    "cfg1_lossless_512": (512, 512, 3, 8, 1, True, 5, 64, None),
    "cfg2_lossy_2048": (2048, 2048, 3, 8, 2, False, 5, 64, 1.0),
}


def make_codestream(name):
    from thousandthchicken_amd import synth
    from tests import imaging
    h, w, c, bits, seed, lossless, lev, cb, bpp = WORKLOADS[name]
    src = synth.image(h, w, c, seed, bits)
    kw = dict(num_resolutions=lev + 1, codeblock_size=(cb, cb), mct=1 if c == 3 else 0)
    if lossless:
        cs = imaging.encode_openjpeg(src if c > 1 else src[..., 0], **kw)
    else:
        rate = (c * bits) / bpp
        cs = imaging.encode_openjpeg(src if c > 1 else src[..., 0], irreversible=True, quality_mode="rates",
                                     quality_layers=[rate], **kw)
    return src, cs


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
    return {"value": pixels * n / el / 1e6, "unit": "Mpixel/s", "cores": 1, "kind": "port",
            "sample": f"{n} full decode(s) of the benchmark codestream by the C oracle, {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="cfg2_lossy_2048", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--bpw-large", type=int, default=0, help="blocks per wave, blocks up to 64x64 (0 auto, 1 wave/block)")
    ap.add_argument("--bpw-small", type=int, default=0, help="blocks per wave, blocks up to 32x32 (0 auto, 1 wave/block)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import thousandthchicken_amd as tca

    src, cs = make_codestream(args.workload)
    info = tca.probe(cs)
    pixels = info.width * info.height
    dec = tca.Decoder()
    dec.set_blocks_per_wave(False, args.bpw_large)
    dec.set_blocks_per_wave(True, args.bpw_small)
    dec.prepare([cs])
    for _ in range(args.warmup):
        dec.run()
    dec.sync()
    # correctness guard on the benchmarked decode
    out = dec.fetch(0)
    h, w, c, bits, seed, lossless, lev, cb, bpp = WORKLOADS[args.workload]
    if lossless and not np.array_equal(out.reshape(src.shape), src):
        raise SystemExit("benchmark decode is not bit-exact")
    if not lossless:
        from oracle import oracle
        ref = oracle.decode(cs)
        dmax = int(np.abs(out.transpose(2, 0, 1).astype(np.int32) - ref).max())
        if dmax > 1:
            raise SystemExit(f"benchmark decode differs from the oracle by {dmax} LSB")

    dec.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dec.run()
    dec.sync()
    el = time.perf_counter() - t0
    ms = el * 1e3 / args.steps

    # per-stage times from HIP events on the decoder's stream
    reps = 10
    acc = np.zeros(5)
    for _ in range(reps):
        dec.run(timed=True)
        dec.sync()
        s = dec.stats()
        acc += [s.fill_ms, s.t1_ms, s.dwt_ms, s.final_ms, s.total_ms]
    acc /= reps
    s = dec.stats()
    t1_bytes = s.codeword_bytes + 4 * s.samples
    t1_gbs = t1_bytes / (acc[1] * 1e-3) / 1e9
    dwt_gbs = (s.dwt_bytes + s.final_bytes) / ((acc[2] + acc[3]) * 1e-3) / 1e9 if acc[2] + acc[3] > 0 else 0.0

    line = {
        "metric": "decoded Mpixel/s on 1 MI355X",
        "value": round(pixels / (ms * 1e-3) / 1e6, 2),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+fp32",
        "data": "synthetic (seeded recipe, coded by OpenJPEG 2.5.4)",
        "config": {"workload": args.workload, "width": info.width, "height": info.height,
                   "components": info.ncomp, "reversible": bool(info.reversible), "levels": info.nlevels,
                   "codeblock": info.cbw, "codestream_bytes": len(cs),
                   "bpp": round(len(cs) * 8 / pixels, 3)},
        "stages_ms": {"fill": round(acc[0], 4), "tier1": round(acc[1], 4), "idwt_levels": round(acc[2], 4),
                      "final_level": round(acc[3], 4), "total": round(acc[4], 4)},
        "roofline": {"bound": "hbm", "kernel": "t1_decode_kernel", "achieved": round(t1_gbs, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(t1_gbs / HBM_PEAK_GBS, 5),
                     "traffic": None, "note": "tier-1 is issue/latency bound; bytes = codewords + 4 B/sample"},
        "roofline_dwt": {"bound": "hbm", "achieved": round(dwt_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(dwt_gbs / HBM_PEAK_GBS, 5)},
        "blocks_per_wave": {"small": s.bpw_small, "large": s.bpw_large},
        "cblks": s.cblks,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(cs, pixels)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
