"""Write the golden fixtures of tests/golden/ from the CPU oracle.

Each fixture is a small code table, an LLR batch and, for every variant, the
posterior / message state after each of the first 5 iterations (reference
layouts post[v*W+w], msg[e*W+w]) plus the final hard decisions.  They pin
later oracle and GPU builds to this oracle's results (regression pins; the
reference itself publishes no vectors).

    python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
from oracle import oracle as O  # noqa: E402

VARIANTS = [O.MS, O.OMS, O.OMS_V2, O.NMS, O.NMS2, O.LOMS, O.LNMS]
ITERS = 5


def toy_codes():
    rng = np.random.default_rng(20261203)
    # (a) regular-ish degree-6 table, 48 variables, 24 checks
    n = 48
    pos_a = np.concatenate([rng.choice(n, 6, replace=False) for _ in range(24)]).astype(np.uint32)
    yield "deg6_n48", n, np.full(24, 6, np.int32), pos_a, -1
    # (b) two degree groups (reference macro form: 8 x deg 5 then 10 x deg 4), Q1 applies
    n = 40
    deg_b = np.array([5] * 8 + [4] * 10, np.int32)
    pos_b = np.concatenate([rng.choice(n, d, replace=False) for d in deg_b]).astype(np.uint32)
    yield "groups5x8_4x10_n40", n, deg_b, pos_b, 8


def main():
    rng = np.random.default_rng(7)
    for name, n, deg, pos, q1 in toy_codes():
        W = 2
        llr = rng.integers(-128, 128, size=4 * W * n).astype(np.int8)
        llr[::7] = rng.integers(-40, 40, size=llr[::7].size)
        out = {"n": n, "deg": deg, "pos": pos, "q1_split": q1 if q1 >= 0 else len(deg), "W": W, "llr": llr}
        table_o = rng.integers(0, 4, size=(ITERS, 9)).astype(np.int8)
        table_n = rng.integers(4, 9, size=(ITERS, 9)).astype(np.int8)
        out["table_loms"], out["table_lnms"] = table_o, table_n
        for v in VARIANTS:
            table = table_o if v == O.LOMS else table_n if v == O.LNMS else None
            dec = O.OracleDecoder(deg, pos, n, W, v, q1_split=out["q1_split"], early_term=0, table=table)
            dec.load(llr)
            posts, msgs = [], []
            for _ in range(ITERS):
                dec.iterate(1)
                posts.append(dec.post.copy())
                msgs.append(dec.msg.copy())
            out[f"post_{v}"] = np.stack(posts)
            out[f"msg_{v}"] = np.stack(msgs)
            out[f"hard_{v}"] = dec.hard()
        np.savez_compressed(HERE / f"{name}.npz", **out)
        print("wrote", HERE / f"{name}.npz")


if __name__ == "__main__":
    main()
