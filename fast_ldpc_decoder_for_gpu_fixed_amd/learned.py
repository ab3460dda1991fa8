"""Learned per-iteration offset tables for OMS on one MI355X: train, quantise,
compare with plain OMS on identical frames.

    python -m fast_ldpc_decoder_for_gpu_fixed_amd.learned [--steps 256 --batch 4096 ...]
    python -m fast_ldpc_decoder_for_gpu_fixed_amd.learned --code ira     # the N=64800 IRA code
    python -m fast_ldpc_decoder_for_gpu_fixed_amd.learned --eval-table tables/loms_3_6_n4000.json  # evaluate only

IRA: the table [iteration][check degree] is trained on a 1/10-length member of
the same ensemble (N=6480, same degree profile: the float shadow's posterior
of N=64800 does not fit LDS) and evaluated on N=64800, with the reference
quantizer trunc(8y) clamped to +-31 (the sat127 quantizer leaves this code in
trapping sets along its degree-2 chain: channel values beyond the 31 message
clamp cannot be overturned).

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

from . import (LEARN_OFFSET, LEARN_SCALE, LNMS, LOMS, OMS, OMS_V2, Code, PoolDecoder, Trainer, code_3_6,
               make_decoder, quantize_offsets, quantize_scales)


def train(code: Code, iters: int, batch: int, steps: int, lr: float, ebn0, seed: int, mode: int = LEARN_OFFSET,
          init: float | None = None, dcols: int | None = None, quant: int = 0):
    dcols = dcols or code.max_deg + 1
    if init is None:
        init = 1.0 if mode == LEARN_OFFSET else 0.75
    tr = Trainer(code, mode, iters, dcols, batch)
    p0 = np.full((iters, dcols), init, np.float32)
    pmax = 8.0 if mode == LEARN_OFFSET else 1.0
    t0 = time.perf_counter()
    params, hist = tr.fit(p0, steps=steps, lr=lr, ebn0=ebn0, seed=seed, first_frame=0, pmin=0.0, pmax=pmax,
                          quant=quant)
    secs = time.perf_counter() - t0
    return params, hist, secs


def evaluate(code: Code, table: np.ndarray, variant: int, iters: int, ebn0_list, frames: int, seed: int,
             words: int = 16384, snr_base: int = 100, quant: int = 0):
    """Decode identical frames (LLR quantizer `quant`) with the plain and learned decoders."""
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
        runs = [(name, lambda d=dec, it=it: d.simulate(eb, frames, it, seed=seed, snr_idx=snr_base + i,
                                                       quant=quant))
                for name, (dec, it) in decs.items()]
        runs += [(name, lambda p=pool: p.simulate(eb, frames, seed=seed, snr_idx=snr_base + i, quant=quant))
                 for name, pool in pools.items()]
        for name, run in runs:
            r = run()
            r["info_gbps"] = r["frames"] * K / (r["decode_ms"] * 1e-3) / 1e9
            row[name] = {k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()}
        out.append(row)
    return out


def evaluate_ira(code: Code, table: np.ndarray, variant: int, max_it: int, ebn0_list, frames: int, seed: int,
                 snr_base: int = 200, fixed_words: int = 2048):
    """Plain OMS (clamp 31) vs the learned table on identical frames (reference
    quantizer): frame-queue decoder with per-frame syndrome stop for both, and
    the drop-in decoder at max_it fixed iterations for plain OMS."""
    K = code.n_info
    out = []
    pools = {"oms_pool": PoolDecoder(code, OMS_V2, max_it), "learned_pool": PoolDecoder(code, variant, max_it,
                                                                                           table=table)}
    fixed = make_decoder(code, fixed_words, OMS_V2, early_term=0)
    for i, eb in enumerate(ebn0_list):
        row = {"ebn0_db": eb}
        for name, pool in pools.items():
            r = pool.simulate(eb, frames, seed=seed, snr_idx=snr_base + i, quant=1)
            r["info_gbps"] = r["frames"] * K / (r["decode_ms"] * 1e-3) / 1e9
            row[name] = {k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()}
        r = fixed.simulate(eb, frames, max_it, seed=seed, snr_idx=snr_base + i, quant=1)
        r["info_gbps"] = r["frames"] * K / (r["decode_ms"] * 1e-3) / 1e9
        row[f"oms{max_it}_fixed"] = {k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()}
        out.append(row)
    return out


def load_table(path: str):
    """(variant, int8 table) from a file written with --save-table."""
    with open(path) as f:
        rec = json.load(f)
    return (LOMS if rec["variant"] == "LOMS" else LNMS), np.asarray(rec["table"], dtype=np.int8)


def table_quantizer(path: str):
    """The LLR quantizer a saved table was trained with (None if not recorded)."""
    with open(path) as f:
        return json.load(f).get("training", {}).get("quantizer")


def ira_codes():
    """(training code N=6480, evaluation code N=64800): info degrees 8 (40%) and
    3 (60%), staircase parity, colour-ordered tables."""
    return (Code.ira(3240, 1296, 8, 3, 11).reorder_by_colour(), Code.ira(32400, 12960, 8, 3, 1).reorder_by_colour())


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--code", choices=["3_6", "ira"], default="3_6")
    ap.add_argument("--mode", choices=["offset", "scale"], default="offset")
    ap.add_argument("--iters", type=int, default=0, help="unrolled iterations (default 10; IRA 20)")
    ap.add_argument("--batch", type=int, default=0, help="frames per step (default 4096; IRA 2048)")
    ap.add_argument("--steps", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-2)
    ap.add_argument("--ebn0-lo", type=float, default=None)
    ap.add_argument("--ebn0-hi", type=float, default=None)
    ap.add_argument("--eval-ebn0", type=float, nargs="+", default=None)
    ap.add_argument("--eval-frames", type=int, default=0, help="default 2^22 (3,6); 2^16 (IRA)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--eval-seed", type=int, default=777)
    ap.add_argument("--save-table", default="", help="write the quantised table and its training record (JSON)")
    ap.add_argument("--no-eval", action="store_true")
    ap.add_argument("--quant", type=int, choices=[0, 1], default=None,
                    help="LLR quantizer: 0 sat127(round(16y/sigma^2)), 1 trunc(8y) clamped to +-31 "
                         "(default: 0 for the (3,6) code, 1 for IRA; with --eval-table, the table's own)")
    ap.add_argument("--eval-table", default="",
                    help="evaluate a table written with --save-table instead of training one")
    ap.add_argument("--drop-last-row", type=int, default=None,
                    help="train one extra unrolled iteration and drop its row (its entry only sharpens the final "
                         "output, so it learns the loss's reward for magnitude rather than decoding; default: on "
                         "for IRA, whose decodes run past the table)")
    a = ap.parse_args(argv)
    ira = a.code == "ira"
    drop = a.drop_last_row if a.drop_last_row is not None else int(ira)
    iters = a.iters or (20 if ira else 10)
    batch = a.batch or (2048 if ira else 4096)
    lo = a.ebn0_lo if a.ebn0_lo is not None else (1.0 if ira else 1.5)
    hi = a.ebn0_hi if a.ebn0_hi is not None else (1.6 if ira else 2.5)
    eval_ebn0 = a.eval_ebn0 or ([1.0, 1.15, 1.3, 1.45, 1.6] if ira else [1.5, 1.75, 2.0, 2.25, 2.5])
    eval_frames = a.eval_frames or ((1 << 16) if ira else (1 << 22))
    quant = a.quant if a.quant is not None else (1 if ira else 0)

    if ira:
        code, eval_code = ira_codes()
        desc = "IRA N=64800 K=32400 (info degrees 8/3, staircase), trained on its N=6480 member"
    else:
        code = eval_code = code_3_6(4000, seed=1)
        desc = "(3,6) N=4000 K=2000, colour-ordered table, seed 1"
    col = code.max_deg
    if a.eval_table:
        variant, table = load_table(a.eval_table)
        if a.quant is None and table_quantizer(a.eval_table) is not None:
            quant = table_quantizer(a.eval_table)
        iters = table.shape[0]
        res = {"code": desc, "table": a.eval_table, "int8_table_deg%d" % col: [int(x) for x in table[:, col]]}
        res["quantizer"] = quant
        res["evaluation"] = (evaluate_ira(eval_code, table, variant, 50, eval_ebn0, eval_frames, a.eval_seed) if ira
                             else evaluate(code, table, variant, iters, eval_ebn0, eval_frames, a.eval_seed,
                                           quant=quant))
        json.dump(res, sys.stdout, indent=1)
        print()
        return
    mode = LEARN_OFFSET if a.mode == "offset" else LEARN_SCALE
    params, hist, secs = train(code, iters + drop, batch, a.steps, a.lr, (lo, hi), a.seed, mode, quant=quant)
    params = params[:iters]
    if mode == LEARN_OFFSET:
        table, variant = quantize_offsets(params), LOMS
    else:
        table, variant = quantize_scales(params), LNMS
    res = {
        "code": desc,
        "training": {"mode": a.mode, "iterations": iters, "batch_frames": batch, "steps": a.steps,
                     "frames": batch * a.steps, "lr": a.lr, "ebn0_range_db": [lo, hi], "quantizer": quant,
                     "trained_iterations": iters + drop, "seed": a.seed, "seconds": round(secs, 2),
                     "loss_first": float(hist[0]),
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
        if ira:
            res["evaluation"] = evaluate_ira(eval_code, table, variant, 50, eval_ebn0, eval_frames, a.eval_seed)
        else:
            res["evaluation"] = evaluate(code, table, variant, iters, eval_ebn0, eval_frames, a.eval_seed,
                                         quant=quant)
    json.dump(res, sys.stdout, indent=1)
    print()


if __name__ == "__main__":
    main()
