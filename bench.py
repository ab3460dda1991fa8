"""Benchmark of the decoding hot path on one MI355X.

One step = one decode of the whole resident batch: 2^16 frames of the
(3,6) N=4000 code, plain OMS (beta = 1 LSB), 20 fixed iterations, no early
stop, LLRs already in HBM.  Prints one JSON line (see README of the task
contract): info Gbit/s, the dominant kernel's HBM roofline and the CPU
baseline (the oracle's serial restatement, timed on the host cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--variant", default="OMS")
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--frames", type=int, default=1 << 16)
    ap.add_argument("--ebn0", type=float, default=2.0)
    ap.add_argument("--words-per-block", type=int, default=0)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--cpu-baseline-words", type=int, default=1024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-json", default=str(ROOT / "profiles" / "hbm_traffic.json"))
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (HBM counter calibration)")
    return ap.parse_args()


def channel_llr(n, frames, ebn0, rate=0.5, seed=1):
    rng = np.random.default_rng(seed)
    sigma = np.sqrt(1.0 / (2 * rate * 10 ** (ebn0 / 10)))
    y = -1.0 + sigma * rng.standard_normal(frames * n, dtype=np.float32)
    return np.clip(np.round(16.0 * y / sigma ** 2), -127, 127).astype(np.int8)


def cpu_baseline(code, llr, n_words, iters, variant):
    """The oracle (scalar C restatement of the reference kernel) on a bounded
    sample of the same workload, all host cores (capped at 16)."""
    from oracle import oracle as O
    deg, pos = code.export()
    N = code.n_vars
    threads = min(16, os.cpu_count() or 1)
    sub = llr[: 4 * n_words * N]
    orc = O.OracleDecoder(deg, pos, N, n_words, variant, q1_split=code.q1_split, early_term=0)
    orc.load(sub)
    t0 = time.perf_counter()
    orc.iterate(iters, threads=threads)
    dt = time.perf_counter() - t0
    frames = 4 * n_words
    return {"value": frames * code.n_info / dt / 1e9, "unit": "info Gbit/s", "cores": threads,
            "kind": "port",
            "sample": f"{frames} frames ({n_words} words) of the same code/input, {iters} iterations, "
                      f"{dt:.2f} s on {threads} host threads"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (frames shard "
              f"across GPUs with no exchange, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import fast_ldpc_decoder_for_gpu_fixed_amd as L

    if args.copy_calibration:
        a, b = L.DeviceBuffer(args.copy_calibration), L.DeviceBuffer(args.copy_calibration)
        L.check(L.load().ldpc_memcpy_d2d(L.C.c_void_p(b.ptr), L.C.c_void_p(a.ptr), args.copy_calibration), "d2d")
        L.synchronize()
        print(json.dumps({"copy_bytes": args.copy_calibration}))
        return

    variant = {"MS": L.MS, "OMS": L.OMS, "OMS_V2": L.OMS_V2, "NMS": L.NMS, "2NMS": L.NMS2}[args.variant.upper()]
    code = L.code_3_6(4000, seed=1)
    N, E, K = code.n_vars, code.n_edges, code.n_info
    frames = args.frames
    W = frames // 4
    llr = channel_llr(N, frames, args.ebn0)
    dec = L.make_decoder(code, W, variant, early_term=0, words_per_block=args.words_per_block,
                         threads=args.threads)
    cfg = dec.config()
    d_in = L.DeviceBuffer(llr.nbytes)
    d_out = L.DeviceBuffer(llr.nbytes)
    x = llr
    if variant == L.OMS_V2:
        from oracle import oracle as O  # layout helper only (input preparation)
        x = O.interleave(llr, N).view(np.int8)
    d_in.upload(x)
    stream = dec.stream

    for _ in range(args.warmup):
        dec.decode_device(d_in.ptr, d_out.ptr, args.iters)
    L.synchronize()
    timer = L.CTimer(stream=stream)
    L.synchronize()
    t0 = time.perf_counter()
    timer.start()
    for _ in range(args.steps):
        dec.decode_device(d_in.ptr, d_out.ptr, args.iters)
    timer.stop()
    ms_total = timer.get_time_ms()
    L.synchronize()
    wall = time.perf_counter() - t0

    ms_step = ms_total / args.steps
    info_gbps = frames * K / (ms_step * 1e-3) / 1e9
    alg_bytes = frames * (args.iters * (2 * E + 2 * N) + 2 * N)
    achieved = alg_bytes / (ms_step * 1e-3) / 1e9

    traffic = None
    tp = Path(args.traffic_json)
    if tp.exists():
        try:
            t = json.loads(tp.read_text())
            if t.get("workload") == f"oms_3_6_n4000_f{frames}_it{args.iters}":
                traffic = t.get("bytes_per_launch")
        except (ValueError, OSError):
            traffic = None

    out = {
        "metric": "info Gbit/s (plain OMS, (3,6) N=4000, 20 it, 2^16 frames, 1 MI355X); HBM GB/s",
        "value": round(info_gbps, 3),
        "unit": "info Gbit/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8x4 (u32 words, saturating int8 lanes)",
        "data": "synthetic: all-zero codeword, BPSK, AWGN at %.1f dB, LLR sat127(round(16y/sigma^2))" % args.ebn0,
        "config": {"workload": "drop-in GPU decoder, (3,6) N=4000 K=2000, %s, %d fixed iterations, "
                               "%d frames resident in HBM" % (L.VARIANT_NAMES[variant], args.iters, frames),
                   "frames": frames, "iterations": args.iters, "variant": L.VARIANT_NAMES[variant],
                   "words_per_block": cfg["words_per_block"], "threads": cfg["threads"],
                   "single_degree_kernel": cfg["uniform_deg"] > 0, "batch": cfg["batch"],
                   "levels_per_iteration": cfg["n_levels"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "wall_s": round(wall, 3),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(code, llr, args.cpu_baseline_words, args.iters, variant)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
