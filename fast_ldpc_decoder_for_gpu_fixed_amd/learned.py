"""Learned per-iteration offset tables for OMS on one MI355X: train, quantise,
compare with plain OMS on identical frames.

    python -m fast_ldpc_decoder_for_gpu_fixed_amd.learned [--steps 256 --batch 4096 ...]

Training runs the float shadow of the layered decoder on the GPU (forward and
backward HIP kernels, Adam on the host) on frames from the device channel with
a per-frame Eb/N0 drawn uniformly in [lo, hi].  The table [iteration][check
degree] is rounded to int8 LSB steps and drives the integer decoder
(LDPC_LOMS), which stays bit-exact against its oracle model.  Evaluation
decodes the same held-out frames (same channel stream) with:
  plain OMS, 20 fixed iterations (the reference point), plain OMS 10 fixed,
  plain OMS with per-word early stop (max 20), learned OMS with per-word early
  stop (max T and max 20), and the persistent frame-queue decoder (per-frame
  stop, slot refill) running plain OMS and learned OMS (max 20),
and reports FER, BER, average iterations per frame and decode throughput.
Prints one JSON document.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import (LEARN_OFFSET, LEARN_SCALE, LNMS, LOMS, OMS, Code, PoolDecoder, Trainer, code_3_6, make_decoder,
               quantize_offsets, quantize_scales)


def train(code: Code, iters: int, batch: int, steps: int, lr: float, ebn0, seed: int, mode: int = LEARN_OFFSET,
          init: float | None = None, dcols: int | None = None):
    dcols = dcols or code.max_deg + 1
    if init is None:
        init = 1.0 if mode == LEARN_OFFSET else 0.75
    tr = Trainer(code, mode, iters, dcols, batch)
    p0 = np.full((iters, dcols), init, np.float32)
    pmax = 8.0 if mode == LEARN_OFFSET else 1.0
    t0 = time.perf_counter()
    params, hist = tr.fit(p0, steps=steps, lr=lr, ebn0=ebn0, seed=seed, first_frame=0, pmin=0.0, pmax=pmax)
    secs = time.perf_counter() - t0
    return params, hist, secs


def evaluate(code: Code, table: np.ndarray, variant: int, iters: int, ebn0_list, frames: int, seed: int,
             words: int = 16384, snr_base: int = 100):
    """Decode identical frames with the plain and learned decoders."""
    K = code.n_info
    out = []
    decs = {
        "oms20": (make_decoder(code, words, OMS, early_term=0), 20),
        "oms10": (make_decoder(code, words, OMS, early_term=0), 10),
        "oms20_early_stop": (make_decoder(code, words, OMS, early_term=1), 20),
        "learned_early_stop": (make_decoder(code, words, variant, early_term=1), iters),
        "learned_fixed": (make_decoder(code, words, variant, early_term=0), iters),
        "learned_early_stop_max20": (make_decoder(code, words, variant, early_term=1), 20),
    }
    for k in ("learned_early_stop", "learned_fixed", "learned_early_stop_max20"):
        decs[k][0].set_table(table)  # rows past the table repeat its last row
    pools = {
        "oms20_pool": PoolDecoder(code, OMS, 20),
        "learned_pool_max20": PoolDecoder(code, variant, 20, table=table),
    }
    for i, eb in enumerate(ebn0_list):
        row = {"ebn0_db": eb}
        runs = [(name, lambda d=dec, it=it: d.simulate(eb, frames, it, seed=seed, snr_idx=snr_base + i))
                for name, (dec, it) in decs.items()]
        runs += [(name, lambda p=pool: p.simulate(eb, frames, seed=seed, snr_idx=snr_base + i))
                 for name, pool in pools.items()]
        for name, run in runs:
            r = run()
            r["info_gbps"] = r["frames"] * K / (r["decode_ms"] * 1e-3) / 1e9
            row[name] = {k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()}
        out.append(row)
    return out


def load_table(path: str):
    """(variant, int8 table) from a file written with --save-table."""
    with open(path) as f:
        rec = json.load(f)
    return (LOMS if rec["variant"] == "LOMS" else LNMS), np.asarray(rec["table"], dtype=np.int8)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--mode", choices=["offset", "scale"], default="offset")
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-2)
    ap.add_argument("--ebn0-lo", type=float, default=1.5)
    ap.add_argument("--ebn0-hi", type=float, default=2.5)
    ap.add_argument("--eval-ebn0", type=float, nargs="+", default=[1.5, 1.75, 2.0, 2.25, 2.5])
    ap.add_argument("--eval-frames", type=int, default=1 << 22)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--eval-seed", type=int, default=777)
    ap.add_argument("--save-table", default="", help="write the quantised table and its training record (JSON)")
    ap.add_argument("--no-eval", action="store_true")
    a = ap.parse_args(argv)

    code = code_3_6(4000, seed=1)
    mode = LEARN_OFFSET if a.mode == "offset" else LEARN_SCALE
    params, hist, secs = train(code, a.iters, a.batch, a.steps, a.lr, (a.ebn0_lo, a.ebn0_hi), a.seed, mode)
    if mode == LEARN_OFFSET:
        table, variant = quantize_offsets(params), LOMS
    else:
        table, variant = quantize_scales(params), LNMS
    col = code.max_deg
    res = {
        "code": "(3,6) N=4000 K=2000, colour-ordered table, seed 1",
        "training": {"mode": a.mode, "iterations": a.iters, "batch_frames": a.batch, "steps": a.steps,
                     "frames": a.batch * a.steps, "lr": a.lr, "ebn0_range_db": [a.ebn0_lo, a.ebn0_hi],
                     "seed": a.seed, "seconds": round(secs, 2), "loss_first": float(hist[0]),
                     "loss_last": float(hist[-1]),
                     "float_table_deg%d" % col: [round(float(x), 4) for x in params[:, col]],
                     "int8_table_deg%d" % col: [int(x) for x in table[:, col]]},
    }
    if a.save_table:
        rec = {"variant": "LOMS" if variant == LOMS else "LNMS", "code": res["code"],
               "table": table.astype(int).tolist(), "training": res["training"]}
        with open(a.save_table, "w") as f:
            json.dump(rec, f, indent=1)
            f.write("\n")
    if not a.no_eval:
        res["evaluation"] = evaluate(code, table, variant, a.iters, a.eval_ebn0, a.eval_frames, a.eval_seed)
    json.dump(res, sys.stdout, indent=1)
    print()


if __name__ == "__main__":
    main()
