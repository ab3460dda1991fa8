"""Benchmark of the decoding hot path on one MI355X.

Default (the headline metric, info Gbit/s at matched FER, learned vs plain OMS):
the (3,6) N=4000 code, 2^20 frames at Eb/N0 = 2.0 dB from the device channel,
resident in HBM.  Two decoders run the same frames:
  * plain  -- the drop-in decoder, plain OMS (beta = 1 LSB), 20 fixed iterations;
  * learned -- the frame-queue decoder (persistent workgroups, per-frame stop,
    slot refill) with the committed learned-offset table, at most 20 iterations.
One step = one decode of the whole batch by the learned decoder; `value` is its
info throughput.  The plain decoder is timed the same way beside it, and both
FERs come from the device error counters on the same frames.

--mode plain: the drop-in decoder alone on 2^16 frames, 20 fixed iterations
(the plain-OMS HBM roofline line); config.variants times all five reference
variants (MS, OMS, OMS_v2, NMS, 2NMS) at 20 and 10 fixed iterations on the same
frames.

--mode ira: the long IRA code (N=64800, K=32400, info degrees 8/3, staircase
parity), 8192 frames with per-frame Eb/N0 uniform in [1.0, 1.6] dB, the
reference quantizer trunc(8y) clamped to +-31, OMS (clamp 31, beta 1), at most
50 iterations: the frame-queue decoder (pair mode: the 4-frame posterior block,
259 KB, does not fit LDS) with the committed learned table and with plain OMS,
against the drop-in decoder at 50 fixed iterations.

--mode mixed: the mixed-code step on two quasi-cyclic codes (12x24 base graph,
Z=81, N=1944, rate 1/2; 4x24, Z=96, N=2304, rate 5/6), 2^17 frames interleaved
50/50, per-frame Eb/N0 uniform in [1.5, 3.5] dB, the reference quantizer
trunc(8y) clamped to +-31, OMS (clamp 31, beta 1), at most 30 iterations with
per-frame stop; each code's kernel autotuned, the step replayed from a hipGraph.

Prints one JSON line: info Gbit/s, the dominant kernel's roofline (algorithmic
bytes / kernel time from HIP events on the kernel's stream) and the CPU
baseline (the oracle's restatement of the reference path on the host cores).
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
TABLE = ROOT / "fast_ldpc_decoder_for_gpu_fixed_amd" / "tables" / "loms_3_6_n4000.json"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", choices=["matched", "plain", "mixed", "ira"], default="matched")
    ap.add_argument("--variant", default="OMS", help="plain mode: MS, OMS, OMS_V2, NMS, 2NMS")
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--frames", type=int, default=0, help="default: 2^20 (matched), 2^16 (plain)")
    ap.add_argument("--ebn0", type=float, default=2.0)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--table", default=str(TABLE))
    ap.add_argument("--words-per-block", type=int, default=0)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--cpu-baseline-words", type=int, default=4096,
                    help="words (4 frames each) of the all-cores CPU baseline sample")
    ap.add_argument("--cpu-single-words", type=int, default=128,
                    help="words of the single-core CPU baseline sample")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed CPU runs after one warm-up (median)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-plain", action="store_true", help="matched mode: time the learned decoder only")
    ap.add_argument("--traffic-json", default=str(ROOT / "profiles" / "hbm_traffic.json"))
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (HBM counter calibration)")
    return ap.parse_args()


def traffic_for(path, key):
    """HBM bytes per launch measured by rocprofv3 counters for this workload."""
    try:
        t = json.loads(Path(path).read_text())
        return t.get("workloads", {}).get(key, {}).get("bytes_per_launch")
    except (ValueError, OSError):
        return None


# The GPU host allots this many CPUs to each GPU; a worker pool larger than the
# share oversubscribes it whatever os.cpu_count() reports.
CPU_SHARE_PER_GPU = 16


def host_cpus():
    """CPU model and the cores this process may use (affinity, cgroup quota,
    the per-GPU share), read at run time."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    logical = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = logical
    quota = None
    try:
        q, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    usable = affinity if quota is None else max(1, min(affinity, int(quota)))
    return {"model": model, "logical_cpus": logical, "affinity_cpus": affinity, "cgroup_cpu_quota": quota,
            "share_per_gpu": CPU_SHARE_PER_GPU, "threads": max(1, min(usable, CPU_SHARE_PER_GPU))}


def cpu_baseline(code, llr, args, iters):
    """The oracle (scalar C restatement of the reference kernel: 4 frames per
    u32, serial layered schedule; plain OMS, fixed iterations) on a bounded
    sample of the same frames: decode only (the load into the word layout is
    outside the timing), one warm-up, median of args.cpu_runs runs, on all
    the cores this process may use, and on one core."""
    from oracle import oracle as O
    from oracle import build as ob
    deg, pos = code.export()
    N = code.n_vars
    host = host_cpus()
    threads = host["threads"]

    def run(words, th):
        orc = O.OracleDecoder(deg, pos, N, words, O.OMS, q1_split=code.q1_split, early_term=0)

        def once():
            orc.load(llr[: 4 * words * N])
            t0 = time.perf_counter()
            orc.iterate(iters, threads=th)
            return time.perf_counter() - t0
        return once

    def timed_decode(words, th):
        once = run(words, th)
        once()  # warm-up
        ts = [once() for _ in range(args.cpu_runs)]
        return float(np.median(ts)), ts

    w_all, w_one = args.cpu_baseline_words, min(args.cpu_single_words, args.cpu_baseline_words)
    t_all, ts_all = timed_decode(w_all, threads)
    t_one, ts_one = timed_decode(w_one, 1)
    return {"value": round(4 * w_all * code.n_info / t_all / 1e9, 5), "unit": "info Gbit/s", "cores": threads,
            "kind": "port",
            "sample": f"{4 * w_all} frames ({w_all} words) of the same code and input, plain OMS, {iters} fixed "
                      f"iterations, decode only; median of {args.cpu_runs} runs after one warm-up, "
                      f"{t_all:.3f} s on {threads} threads",
            "runs_s": [round(t, 4) for t in ts_all],
            "host": host, "compiler": ob.compiler_version(), "flags": " ".join(ob.CFLAGS),
            "single_core": {"value": round(4 * w_one * code.n_info / t_one / 1e9, 6), "unit": "info Gbit/s",
                            "sample": f"first {4 * w_one} frames, median of {args.cpu_runs} runs after one "
                                      f"warm-up, {t_one:.3f} s on 1 thread",
                            "runs_s": [round(t, 4) for t in ts_one]}}


def timed(L, stream, steps, warmup, fn):
    """ms per step of fn() over `steps` calls, HIP events on `stream`."""
    for _ in range(warmup):
        fn()
    L.synchronize()
    timer = L.CTimer(stream=stream)
    timer.start()
    for _ in range(steps):
        fn()
    timer.stop()
    ms = timer.get_time_ms()
    L.synchronize()
    return ms / steps


def counts(L, dec, d_hard, frames, n_info):
    d_cnt = L.DeviceBuffer(24)
    dec.count_errors(d_hard, frames, n_info, d_cnt.ptr, stream=dec.stream)
    L.synchronize()
    c = d_cnt.download(np.zeros(3, np.uint64))
    return {"frame_errors": int(c[1]), "bit_errors": int(c[0]), "fer": float(c[1]) / frames,
            "ber": float(c[0]) / (frames * n_info)}


def roofline(alg_bytes, ms, traffic):
    """HBM roofline of the dominant kernel: algorithmic bytes per launch over
    its HIP-event time.  With the posterior held in LDS the kernels issue
    fewer HBM bytes than the algorithmic figure and their limiter is the VALU
    (rocprofv3 counters in profiles/); frac_measured is the profiled HBM bytes
    per launch over the same time."""
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "algorithmic_bytes": int(alg_bytes),
           "limiter": "valu"}
    if traffic:
        out["frac_measured"] = round(traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    return out


HBM_GBPS_METHOD = ("rocprofv3 FETCH_SIZE + WRITE_SIZE bytes per launch of this workload (separate counter runs, "
                   "calibrated on a device copy; profiles/hbm_traffic.json) over this run's kernel time")


def run_matched(L, args, t_start):
    from fast_ldpc_decoder_for_gpu_fixed_amd.learned import load_table
    code = L.code_3_6(4000, seed=1)
    N, E, K = code.n_vars, code.n_edges, code.n_info
    frames = args.frames or (1 << 20)
    if frames % 4:
        raise SystemExit("frames must be a multiple of 4")
    variant, table = load_table(args.table)
    sigma = L.channel_sigma(args.ebn0, code.rate)
    d_llr = L.DeviceBuffer(frames * N)
    L.channel_generate(d_llr.ptr, N, frames, args.seed, 0, sigma)  # frame-major int8, Philox AWGN
    L.synchronize()

    # learned: frame-queue decoder, per-frame stop, max `iters`; the step is
    # the captured hipGraph (queue reset + frame-queue kernel), eager beside it
    probe = L.make_decoder(code, 4, L.OMS)
    d_hard_l, d_it = L.DeviceBuffer(frames * N), L.DeviceBuffer(4 * frames)

    def queue_leg(v, tab):
        pool = L.PoolDecoder(code, v, args.iters, table=tab, words_per_block=args.words_per_block,
                             threads=args.threads)
        ms_eager = timed(L, pool.stream, args.steps, args.warmup,
                         lambda: pool.decode_device(d_llr.ptr, frames, d_hard_l.ptr, d_it.ptr))
        ms = timed(L, pool.stream, args.steps, args.warmup,
                   lambda: pool.decode_device(d_llr.ptr, frames, d_hard_l.ptr, d_it.ptr, graph=True))
        its = d_it.download(np.empty(frames, np.int32))
        r = counts(L, probe, d_hard_l.ptr, frames, K)
        it_sum = int(its.sum(dtype=np.int64))
        r.update({"info_gbps": round(frames * K / (ms * 1e-3) / 1e9, 3), "ms_per_step": round(ms, 4),
                  "graph_ms_per_step": round(ms, 4), "eager_ms_per_step": round(ms_eager, 4),
                  "avg_iterations": round(it_sum / frames, 4), "max_iterations": args.iters,
                  "decoder": "frame-queue (persistent, per-frame stop, refill), step replayed from a hipGraph"})
        r.update(pool.config())
        return r, ms, it_sum

    learned, ms_l, iters_sum = queue_leg(variant, table)
    learned["table_deg6"] = [int(x) for x in table[:, min(6, table.shape[1] - 1)]]
    alg_l = iters_sum * (2 * E + 2 * N) + frames * 2 * N
    gbps_l = frames * K / (ms_l * 1e-3) / 1e9
    # plain OMS in the same frame queue on the same frames: what the table alone gains
    plain_queue = None
    if not args.skip_plain:
        plain_queue, _, _ = queue_leg(L.OMS, None)
    del d_hard_l

    plain = None
    if not args.skip_plain:
        dec = L.make_decoder(code, frames // 4, L.OMS, early_term=0)
        d_hard_p = L.DeviceBuffer(frames * N)
        ms_p = timed(L, dec.stream, args.steps, args.warmup,
                     lambda: dec.decode_device(d_llr.ptr, d_hard_p.ptr, args.iters))
        plain = counts(L, dec, d_hard_p.ptr, frames, K)
        alg_p = frames * (args.iters * (2 * E + 2 * N) + 2 * N)
        plain.update({"info_gbps": round(frames * K / (ms_p * 1e-3) / 1e9, 3), "ms_per_step": round(ms_p, 4),
                      "avg_iterations": args.iters, "decoder": "drop-in, fixed iterations"})
        plain["roofline"] = roofline(alg_p, ms_p, traffic_for(args.traffic_json,
                                                             f"oms_3_6_n4000_f{frames}_it{args.iters}"))
        cfg = dec.config()
        plain.update({"words_per_block": cfg["words_per_block"], "threads": cfg["threads"],
                      "single_degree_kernel": cfg["uniform_deg"] > 0, "batch": cfg["batch"]})
        del dec, d_hard_p

    key = f"loms_pool_3_6_n4000_f{frames}_ebn0_{args.ebn0:.2f}_max{args.iters}"
    out = {
        "metric": "info Gbit/s at matched FER, learned vs plain OMS, 1 MI355X; HBM GB/s",
        "value": round(gbps_l, 3),
        "unit": "info Gbit/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_l, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8x4 (u32 words, saturating int8 lanes)",
        "data": "synthetic: all-zero codeword, BPSK bit0 -> -1, device Philox AWGN at %.2f dB (seed %d), "
                "LLR sat127(round(16y/sigma^2))" % (args.ebn0, args.seed),
        "config": {"workload": "learned-offset OMS vs plain 20-iteration OMS, (3,6) N=4000 K=2000, "
                               "%d frames resident in HBM, Eb/N0 %.2f dB" % (frames, args.ebn0),
                   "frames": frames, "ebn0_db": args.ebn0, "parallelism": "single GPU",
                   "graph_ms_per_step": learned["graph_ms_per_step"],
                   "eager_ms_per_step": learned["eager_ms_per_step"],
                   "learned": learned, "plain": plain, "plain_frame_queue": plain_queue},
        "roofline": roofline(alg_l, ms_l, traffic_for(args.traffic_json, key)),
    }
    tr = out["roofline"]["traffic"]
    if tr:  # the metric's "HBM GB/s"
        out["config"]["hbm_gbps"] = round(tr / (ms_l * 1e-3) / 1e9, 1)
        out["config"]["hbm_gbps_method"] = HBM_GBPS_METHOD
    if plain:
        # learned frame queue vs plain drop-in at 20 fixed iterations: the
        # per-frame stop and the table together
        out["config"]["speedup_vs_plain"] = round(gbps_l / plain["info_gbps"], 3)
        out["config"]["fer_matched"] = learned["fer"] <= plain["fer"]
    if plain_queue:
        # the table alone: learned vs plain OMS in the same frame queue
        out["config"]["speedup_vs_plain_frame_queue"] = round(gbps_l / plain_queue["info_gbps"], 3)
        out["config"]["speedup_parts"] = ("speedup_vs_plain = per-frame stop x table; "
                                          "speedup_vs_plain_frame_queue = the table alone")
    if not args.no_cpu_baseline:
        n = 4 * args.cpu_baseline_words * N
        llr = d_llr.download(np.empty(n, np.int8))
        out["cpu_baseline"] = cpu_baseline(code, llr, args, args.iters)
    out["wall_s"] = round(time.perf_counter() - t_start, 3)
    return out


def run_plain(L, args, t_start):
    variant = {"MS": L.MS, "OMS": L.OMS, "OMS_V2": L.OMS_V2, "NMS": L.NMS, "2NMS": L.NMS2}[args.variant.upper()]
    code = L.code_3_6(4000, seed=1)
    N, E, K = code.n_vars, code.n_edges, code.n_info
    frames = args.frames or (1 << 16)
    W = frames // 4
    dec = L.make_decoder(code, W, variant, early_term=0, words_per_block=args.words_per_block,
                         threads=args.threads)
    cfg = dec.config()
    sigma = L.channel_sigma(args.ebn0, code.rate)
    d_in = L.DeviceBuffer(frames * N)
    # OMS_v2 takes its input in the device word layout (the reference's xMS convention)
    L.channel_generate(d_in.ptr, N, frames, args.seed, 0, sigma, layout=1 if variant == L.OMS_V2 else 0)
    d_out = L.DeviceBuffer(frames * N)
    ms = timed(L, dec.stream, args.steps, args.warmup, lambda: dec.decode_device(d_in.ptr, d_out.ptr, args.iters))
    info_gbps = frames * K / (ms * 1e-3) / 1e9
    alg = frames * (args.iters * (2 * E + 2 * N) + 2 * N)
    # the drop-in's host-buffer decode (pinned in/out): PCIe copies included, never `value`
    pin_in, pin_out = L.PinnedBuffer(frames * N), L.PinnedBuffer(frames * N)
    d_in.download(pin_in.array)
    dec.decode(pin_in.array, pin_out.array, args.iters)
    t0 = time.perf_counter()
    for _ in range(3):
        dec.decode(pin_in.array, pin_out.array, args.iters)
    ms_pcie = (time.perf_counter() - t0) * 1e3 / 3
    same = bool(np.array_equal(pin_out.array, d_out.download(np.empty(frames * N, np.uint8))))
    out = {
        "metric": "info Gbit/s (plain %s, (3,6) N=4000, %d it, %d frames, 1 MI355X); HBM GB/s"
                  % (L.VARIANT_NAMES[variant], args.iters, frames),
        "value": round(info_gbps, 3),
        "unit": "info Gbit/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8x4 (u32 words, saturating int8 lanes)",
        "data": "synthetic: all-zero codeword, BPSK, device Philox AWGN at %.1f dB, LLR sat127(round(16y/sigma^2))"
                % args.ebn0,
        "config": {"workload": "drop-in GPU decoder, (3,6) N=4000 K=2000, %s, %d fixed iterations, "
                               "%d frames resident in HBM" % (L.VARIANT_NAMES[variant], args.iters, frames),
                   "frames": frames, "iterations": args.iters, "variant": L.VARIANT_NAMES[variant],
                   "words_per_block": cfg["words_per_block"], "threads": cfg["threads"],
                   "single_degree_kernel": cfg["uniform_deg"] > 0, "batch": cfg["batch"],
                   "levels_per_iteration": cfg["n_levels"], "parallelism": "single GPU",
                   "pcie_inclusive": {"ms_per_decode": round(ms_pcie, 3),
                                      "info_gbps": round(frames * K / (ms_pcie * 1e-3) / 1e9, 3),
                                      "buffers": "pinned host in/out, H2D + decode + D2H in 8 overlapped chunks, "
                                                 "synchronous call",
                                      "decisions_equal_device_path": same}},
        "roofline": roofline(alg, ms, traffic_for(args.traffic_json, f"oms_3_6_n4000_f{frames}_it{args.iters}")
                             if variant == L.OMS else None),
    }
    tr = out["roofline"]["traffic"]
    if tr:
        out["config"]["hbm_gbps"] = round(tr / (ms * 1e-3) / 1e9, 1)
        out["config"]["hbm_gbps_method"] = HBM_GBPS_METHOD
    # the five reference variants at 10 and 20 fixed iterations on the same
    # frames (the drop-in workload's full set), each timed the same way; the
    # OMS_v2 decoder reads the word layout, the others frame-major frames
    if variant == L.OMS_V2:
        d_words, d_frames = d_in, L.DeviceBuffer(frames * N)
        L.channel_generate(d_frames.ptr, N, frames, args.seed, 0, sigma, layout=0)
    else:
        d_frames, d_words = d_in, L.DeviceBuffer(frames * N)
        L.interleave_device(d_in.ptr, d_words.ptr, N, W)
    L.synchronize()
    table = {}
    for name, v in (("MS", L.MS), ("OMS", L.OMS), ("OMS_v2", L.OMS_V2), ("NMS", L.NMS), ("2NMS", L.NMS2)):
        dv = L.make_decoder(code, W, v, early_term=0, words_per_block=args.words_per_block, threads=args.threads)
        src = d_words.ptr if v == L.OMS_V2 else d_frames.ptr
        row = {}
        for it in (20, 10):
            t = timed(L, dv.stream, min(args.steps, 20), 1, lambda: dv.decode_device(src, d_out.ptr, it))
            c = counts(L, dv, d_out.ptr, frames, K)
            row[f"it{it}"] = {"ms_per_decode": round(t, 4), "info_gbps": round(frames * K / (t * 1e-3) / 1e9, 3),
                              "fer": c["fer"], "ber": c["ber"]}
        table[name] = row
        del dv
    out["config"]["variants"] = table
    if not args.no_cpu_baseline and variant == L.OMS:
        n = 4 * args.cpu_baseline_words * N
        llr = d_in.download(np.empty(n, np.int8))
        out["cpu_baseline"] = cpu_baseline(code, llr, args, args.iters)
    out["wall_s"] = round(time.perf_counter() - t_start, 3)
    return out


def run_ira(L, args, t_start):
    from fast_ldpc_decoder_for_gpu_fixed_amd.learned import ira_codes, load_table
    code = ira_codes()[1]
    N, E, K = code.n_vars, code.n_edges, code.n_info
    frames = args.frames or 8192
    max_it = 50
    variant, table = load_table(str(ROOT / "fast_ldpc_decoder_for_gpu_fixed_amd" / "tables" / "loms_ira_n64800.json"))
    d_llr = L.DeviceBuffer(frames * N)
    L.channel_generate_range(d_llr.ptr, N, frames, args.seed, 1.0, 1.6, code.rate, quant=1)
    L.synchronize()
    probe = L.make_decoder(code, 4, L.OMS_V2)
    legs = {}
    d_hard, d_it = L.DeviceBuffer(frames * N), L.DeviceBuffer(4 * frames)
    # the frame queues take the noisiest frames first (ldpc_pool_set_order); the learned decoder is also timed
    # in input order, the queue order of earlier measurements
    lrf, inp = L.ORDER_LEAST_RELIABLE_FIRST, L.ORDER_INPUT
    for name, v, tab, order in (("learned_pool", variant, table, lrf), ("oms_pool", L.OMS_V2, None, lrf),
                                ("learned_pool_input_order", variant, table, inp)):
        pool = L.PoolDecoder(code, v, max_it, table=tab, words_per_block=args.words_per_block,
                             threads=args.threads or 1024, order=order)
        ms = timed(L, pool.stream, args.steps, args.warmup,
                   lambda: pool.decode_device(d_llr.ptr, frames, d_hard.ptr, d_it.ptr))
        its = d_it.download(np.empty(frames, np.int32)).astype(np.int64)
        r = counts(L, probe, d_hard.ptr, frames, K)
        alg = int(its.sum()) * (2 * E + 2 * N) + frames * 2 * N
        r.update({"info_gbps": round(frames * K / (ms * 1e-3) / 1e9, 3), "ms_per_step": round(ms, 4),
                  "avg_iterations": round(float(its.mean()), 4),
                  "order": "least reliable first" if order == lrf else "input",
                  "roofline": roofline(alg, ms, traffic_for(args.traffic_json, f"ira_pool_n64800_f{frames}")
                                       if name == "learned_pool" else None)})
        r.update(pool.config())
        legs[name] = r
        del pool
    dec = L.make_decoder(code, frames // 4, L.OMS_V2, early_term=0)
    d_words = L.DeviceBuffer(frames * N)  # OMS_v2 takes the device word layout (the reference's xMS convention)
    L.interleave_device(d_llr.ptr, d_words.ptr, N, frames // 4)
    L.synchronize()
    ms_f = timed(L, dec.stream, max(1, args.steps // 5), 1,
                 lambda: dec.decode_device(d_words.ptr, d_hard.ptr, max_it))
    r = counts(L, dec, d_hard.ptr, frames, K)
    r.update({"info_gbps": round(frames * K / (ms_f * 1e-3) / 1e9, 3), "ms_per_step": round(ms_f, 4),
              "avg_iterations": max_it, "decoder": "drop-in, fixed iterations (long frames: pair-mode frame queue, no stop test)",
              "roofline": roofline(frames * (max_it * (2 * E + 2 * N) + 2 * N), ms_f, None)})
    legs["oms50_fixed"] = r
    lead = legs["learned_pool"]
    out = {
        "metric": "info Gbit/s, IRA N=64800 frame-queue decode (pair mode), learned vs plain OMS, 1 MI355X",
        "value": lead["info_gbps"], "unit": "info Gbit/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": lead["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8x2 (2 frames per packed-16 lane pair)",
        "data": "synthetic: all-zero codeword, BPSK, device Philox AWGN, per-frame Eb/N0 U[1.0, 1.6] dB, "
                "LLR trunc(8y) clamped to +-31",
        "config": {"workload": "IRA N=64800 K=32400 (info degrees 8/3, staircase), %d frames, OMS clamp 31, max %d "
                               "iterations, per-frame syndrome stop" % (frames, max_it),
                   "edges": E, "levels": code.n_levels, "legs": legs,
                   "speedup_vs_fixed": round(lead["info_gbps"] / legs["oms50_fixed"]["info_gbps"], 3),
                   "parallelism": "single GPU"},
        "roofline": lead["roofline"],
    }
    out["wall_s"] = round(time.perf_counter() - t_start, 3)
    return out


def run_mixed(L, args, t_start):
    codes = [L.Code.qc(12, 24, 81, 0.35, seed=1), L.Code.qc(4, 24, 96, 0.35, seed=2)]
    names = ["qc_12x24_z81", "qc_4x24_z96"]
    frames = args.frames or (1 << 17)
    max_it = 30
    variant = L.OMS_V2  # OMS with the 31 clamp on every check (these codes have many check degrees)
    # frame f belongs to code f & 1; each code's frames are contiguous in its own region
    cid = (np.arange(frames) & 1).astype(np.uint8)
    n_c = [(frames + 1 - k) // 2 for k in range(2)]
    base = [0, n_c[0] * codes[0].n_vars]
    off = np.empty(frames, np.uint64)
    for k in range(2):
        off[k::2] = base[k] + np.arange(n_c[k], dtype=np.uint64) * codes[k].n_vars
    total = base[1] + n_c[1] * codes[1].n_vars
    d_in, d_out = L.DeviceBuffer(total), L.DeviceBuffer(total)
    d_off, d_cid, d_it = L.DeviceBuffer(off.nbytes), L.DeviceBuffer(frames), L.DeviceBuffer(4 * frames)
    d_off.upload(off)
    d_cid.upload(cid)
    for k, c in enumerate(codes):
        L.channel_generate_range(d_in.ptr + base[k], c.n_vars, n_c[k], args.seed + k, 1.5, 3.5, c.rate, quant=1)
    L.synchronize()
    t_tune = time.perf_counter()
    tuned = [L.pool_autotune(c, variant, max_it, ebn0_db=2.5, sample_frames=65536, quant=1) for c in codes]
    t_tune = time.perf_counter() - t_tune
    mixed = L.MixedDecoder(codes, variant, max_it, tunings=tuned, max_frames=frames)
    step = lambda graph: mixed.decode_device(d_in.ptr, d_off.ptr, d_cid.ptr, frames, d_out.ptr, d_it.ptr,
                                             graph=graph)
    ms_eager = timed(L, 0, args.steps, args.warmup, lambda: step(False))
    ms = timed(L, 0, args.steps, args.warmup, lambda: step(True))
    its = d_it.download(np.empty(frames, np.int32))
    # each code's frames decoded alone (its own frame queue, same tuning, the
    # contiguous region of its frames): the per-code rate beside the shared step
    alone = []
    for k, c in enumerate(codes):
        t = tuned[k]
        pool = L.PoolDecoder(c, variant, max_it, words_per_block=t["words_per_block"], threads=t["threads"],
                             uniform=t["uniform"], batch=t["batch"], pairs=t["pairs"])
        d_o, d_i = L.DeviceBuffer(n_c[k] * c.n_vars), L.DeviceBuffer(4 * n_c[k])
        ms_k = timed(L, pool.stream, max(5, args.steps // 4), args.warmup,
                     lambda: pool.decode_device(d_in.ptr + base[k], n_c[k], d_o.ptr, d_i.ptr))
        alone.append(ms_k)
        del pool
    per_code, alg, info = [], 0, 0
    for k, c in enumerate(codes):
        ik = its[k::2].astype(np.int64)
        r = counts(L, L.make_decoder(c, 4, L.OMS_V2), d_out.ptr + base[k], n_c[k], c.n_info)
        r.update({"code": names[k], "n": c.n_vars, "k": c.n_info, "edges": c.n_edges, "max_check_degree": c.max_deg,
                  "levels": c.n_levels, "frames": n_c[k], "avg_iterations": round(float(ik.mean()), 4),
                  "iterations_hist": np.bincount(ik, minlength=max_it + 1).tolist(),
                  "tuning": {kk: tuned[k][kk] for kk in ("words_per_block", "threads", "uniform", "batch", "pairs")},
                  "tune_ms_per_65536_frames": round(tuned[k]["ms"], 4),
                  "alone_ms": round(alone[k], 4),
                  "alone_info_gbps": round(n_c[k] * c.n_info / (alone[k] * 1e-3) / 1e9, 3)})
        per_code.append(r)
        alg += int(ik.sum()) * (2 * c.n_edges + 2 * c.n_vars) + n_c[k] * 2 * c.n_vars
        info += n_c[k] * c.n_info
    gbps = info / (ms * 1e-3) / 1e9
    out = {
        "metric": "info Gbit/s, mixed-code frame-queue decode step (2 QC codes), 1 MI355X; HBM GB/s",
        "value": round(gbps, 3), "unit": "info Gbit/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8x4 (u32 words, saturating int8 lanes)",
        "data": "synthetic: all-zero codewords, BPSK, device Philox AWGN, per-frame Eb/N0 U[1.5, 3.5] dB, "
                "LLR trunc(8y) clamped to +-31",
        "config": {"workload": "mixed QC 12x24 Z=81 + 4x24 Z=96, %d frames interleaved 50/50, OMS clamp 31, "
                               "max %d iterations, per-frame stop, hipGraph step" % (frames, max_it),
                   "frames": frames, "codes": per_code, "eager_ms_per_step": round(ms_eager, 4),
                   "eager_info_gbps": round(info / (ms_eager * 1e-3) / 1e9, 3),
                   "autotune_seconds": round(t_tune, 2), "parallelism": "single GPU"},
        "roofline": roofline(alg, ms, traffic_for(args.traffic_json, f"mixed_qc_12x24_z81_4x24_z96_f{frames}")),
    }
    tr = out["roofline"]["traffic"]
    if tr:  # the metric's "HBM GB/s": the step's kernels together
        out["config"]["hbm_gbps"] = round(tr / (ms * 1e-3) / 1e9, 1)
        out["config"]["hbm_gbps_method"] = HBM_GBPS_METHOD
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        from oracle import build as ob
        host = host_cpus()
        threads = host["threads"]
        nf = 1024
        samples, bits = [], 0
        raw = d_in.download(np.empty(base[1] + nf * codes[1].n_vars, np.int8))
        for k, c in enumerate(codes):
            deg, pos = c.export()
            llr = raw[base[k]:base[k] + nf * c.n_vars].reshape(nf, c.n_vars)
            samples.append((deg, pos, c.n_vars, llr))
            bits += nf * c.n_info

        def once():
            t0 = time.perf_counter()
            for deg, pos, n, llr in samples:
                O.decode_frames(deg, pos, n, O.OMS_V2, llr, max_it, threads=threads)
            return time.perf_counter() - t0
        once()  # warm-up
        ts = [once() for _ in range(args.cpu_runs)]
        t_cpu = float(np.median(ts))
        out["cpu_baseline"] = {"value": round(bits / t_cpu / 1e9, 5), "unit": "info Gbit/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{nf} frames of each code from the same input, per-frame stop, max {max_it}; "
                                         f"median of {args.cpu_runs} runs after one warm-up, {t_cpu:.3f} s on "
                                         f"{threads} threads",
                               "runs_s": [round(t, 4) for t in ts], "host": host,
                               "compiler": ob.compiler_version(), "flags": " ".join(ob.CFLAGS)}
    out["wall_s"] = round(time.perf_counter() - t_start, 3)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (frames shard across GPUs "
              f"with no exchange; the multi-GPU layout is a design note only)", file=sys.stderr)
        sys.exit(2)
    import fast_ldpc_decoder_for_gpu_fixed_amd as L

    if args.copy_calibration:
        a, b = L.DeviceBuffer(args.copy_calibration), L.DeviceBuffer(args.copy_calibration)
        L.check(L.load().ldpc_memcpy_d2d(L.C.c_void_p(b.ptr), L.C.c_void_p(a.ptr), args.copy_calibration), "d2d")
        L.synchronize()
        print(json.dumps({"copy_bytes": args.copy_calibration}))
        return
    t0 = time.perf_counter()
    run = {"matched": run_matched, "plain": run_plain, "mixed": run_mixed, "ira": run_ira}[args.mode]
    out = run(L, args, t0)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
