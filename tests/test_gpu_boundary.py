"""GPU check core at the boundary values of the int8 arithmetic.

The oracle's check update is brute-forced against a textbook statement over
the boundary values (tests/test_oracle_decoder.py); this file drives the same
values through every GPU path that runs the packed-16 check core, and
compares each with the oracle bit for bit:

  * k_decode: generic degree classes, the single-degree kernel with 1, 2 and
    4 items in flight, the global-memory posterior; posterior and messages
    after every iteration (state API);
  * k_pool: 4-frame words (generic and single-degree with 1, 2, 4 items in
    flight) and 2-frame pairs (rows layout for degrees <= 8, the generic pair
    gather above; two pairs per workgroup and one); decisions and iteration
    counts per frame.

Semantics at stake (src/decoder_ms/cuda/simd_functions.h:89-131,
src/decoder_oms/cuda/CUDA_OMS_SIMD.cu:166-186): |-128| = 128 never becomes a
minimum, the tie select (|t| == min1), saturation at -128/127, zero counted
as non-positive, learned offsets up to 127 and scales 0 and 8.

Codes: layer A lists disjoint checks that cover every variable once, in
order, so the channel values of a frame enumerate A's check inputs (all
degree-2 and degree-3 combinations of the boundary values; ties, all-equal
and extreme tuples for the larger degrees); layer B covers a seeded
permutation of the variables with the same degrees, so it reads the
posteriors A left (saturated sums included).  Three iterations: the second
and third subtract the previous messages."""
import itertools
import os
import sys

import numpy as np
import pytest

import fast_ldpc_decoder_for_gpu_fixed_amd as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
EDGE_VALUES = [-128, -127, -100, -33, -32, -31, -30, -2, -1, 0, 1, 2, 30, 31, 32, 33, 96, 126, 127]
ITERS = 3


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def tuples_for(d, count, rng):
    """Input tuples for degree-d checks: every combination for d <= 3,
    else random boundary tuples with forced ties and extreme patterns."""
    if d <= 3:
        return [list(t) for t in itertools.product(EDGE_VALUES, repeat=d)]
    out = [[v] * d for v in (-128, 127, 0, -1, 1, -31, 31, 32)]  # all equal
    out += [[-128] * (d - 1) + [127], [127] * (d - 1) + [-128], [0] * (d - 1) + [-128]]
    while len(out) < count:
        t = list(rng.choice(EDGE_VALUES, d))
        r = rng.random()
        if r < 0.4:  # a tie on the minimum magnitude
            mags = [abs(x) if x != -128 else 128 for x in t]
            i = int(np.argmin(mags))
            j = (i + 1 + int(rng.integers(d - 1))) % d
            t[j] = t[i] if rng.random() < 0.5 else -t[i] if t[i] != -128 else 127
        elif r < 0.55:  # three-way tie
            v = int(rng.choice(EDGE_VALUES))
            for j in rng.choice(d, 3, replace=False):
                t[j] = v
        out.append([int(x) for x in t])
    return out


class BoundaryCode:
    def __init__(self, name, degs, seed):
        self.name = name
        degs = [int(x) for x in degs]
        n = sum(degs)
        rng = np.random.default_rng(seed)
        perm = rng.permutation(n)
        degs_b = [int(x) for x in rng.permutation(degs)]
        deg = np.array(degs + degs_b, np.int32)
        pos = np.concatenate([np.arange(n), perm]).astype(np.uint32)
        self.code = L.Code.from_table(n, deg, pos)
        self.deg, self.pos = self.code.export()
        self.n = n
        self.degs_a = degs
        self.starts = np.concatenate([[0], np.cumsum(degs)[:-1]])
        # frame values: A-check i of frame f takes tuple f * len(A) + i of its degree's list
        lists = {d: tuples_for(d, 1024, rng) for d in set(degs)}
        need = max(-(-len(lists[d]) // degs.count(d)) for d in lists)
        F = max(16, -(-need // 16) * 16)
        llr = np.empty((F, n), np.int8)
        cursor = {d: 0 for d in lists}
        for f in range(F):
            for i, d in enumerate(degs):
                lst = lists[d]
                k = cursor[d]
                cursor[d] += 1
                t = lst[k] if k < len(lst) else list(rng.choice(EDGE_VALUES, d))
                llr[f, self.starts[i]:self.starts[i] + d] = t
        self.llr = llr
        self.F = F
        assert all(cursor[d] >= len(lists[d]) for d in lists), "every enumerated tuple is used"


_CODES = {}


def codes():
    if not _CODES:
        spec = [("d2", [2] * 64), ("d3", [3] * 64), ("d6", [6] * 64), ("d7", [7] * 48), ("d8", [8] * 48),
                ("d12", [12] * 24), ("mix8", [2, 3, 6, 7, 8] * 12), ("mix10", [3, 9, 10] * 12),
                ("mix16", [2, 6, 12, 16] * 8), ("d20", [20] * 12)]
        for k, (name, degs) in enumerate(spec):
            _CODES[name] = BoundaryCode(name, degs, seed=100 + k)
    return _CODES


def const_table(v, iters=ITERS + 2):
    return np.full((iters, 33), v, np.int8)


def rand_table(variant, seed, iters=ITERS + 2):
    rng = np.random.default_rng(seed)
    hi = 127 if variant == L.LOMS else 8
    t = rng.integers(0, hi + 1, size=(iters, 33)).astype(np.int8)
    t[:, ::5] = hi  # the extremes on some degrees
    t[:, 1::7] = 0
    return t


# (variant, table) cases: the five reference variants, then learned offsets
# beta in {0, 31, 127} and scales alpha in {0, 8} (and mixed tables)
CASES = [(L.MS, None), (L.OMS, None), (L.OMS_V2, None), (L.NMS, None), (L.NMS2, None),
         (L.LOMS, const_table(0)), (L.LOMS, const_table(31)), (L.LOMS, const_table(127)),
         (L.LOMS, rand_table(L.LOMS, 1)), (L.LNMS, const_table(0)), (L.LNMS, const_table(8)),
         (L.LNMS, rand_table(L.LNMS, 2))]


def case_name(v, t):
    return L.VARIANT_NAMES[v] + ("" if t is None else f"[{int(t.min())}..{int(t.max())}]")


DECODE_PATHS = {
    "generic": dict(uniform=0, words_per_block=2, threads=256),
    "uniform_g1": dict(uniform=1, batch=1, words_per_block=2, threads=256),
    "uniform_g2": dict(uniform=1, batch=2, words_per_block=2, threads=256),
    "uniform_g4": dict(uniform=1, batch=4, words_per_block=2, threads=256),
    "global_post": dict(post_in_lds=0, words_per_block=4, threads=256),
}


@pytest.mark.parametrize("path", list(DECODE_PATHS))
def test_decoder_state_at_boundary_values(path, gpu_available):
    tuning = DECODE_PATHS[path]
    ran = 0
    for bc in codes().values():
        if tuning.get("uniform") == 1 and not (len(set(bc.deg.tolist())) == 1 and 6 <= bc.deg[0] <= 8):
            continue
        W = bc.F // 4
        for variant, table in CASES:
            say(f"k_decode {path}: code {bc.name} N={bc.n} W={W} {case_name(variant, table)}")
            dec = L.make_decoder(bc.code, W, variant, early_term=0, **tuning)
            if tuning.get("uniform") == 1:
                assert dec.config()["uniform_deg"] == bc.deg[0]
            if table is not None:
                dec.set_table(table)
            orc = O.OracleDecoder(bc.deg, bc.pos, bc.n, W, variant, q1_split=bc.code.q1_split, early_term=0,
                                  table=table)
            orc.load(bc.llr)
            dec.load(O.interleave(bc.llr, bc.n).view(np.int8) if variant == L.OMS_V2 else bc.llr)
            for it in range(1, ITERS + 1):
                dec.iterate(1)
                orc.iterate(1, threads=THREADS)
                post, msg = dec.get_state()
                np.testing.assert_array_equal(post, orc.post, err_msg=f"{bc.name}: posterior after iteration {it}")
                np.testing.assert_array_equal(msg, orc.msg, err_msg=f"{bc.name}: messages after iteration {it}")
            np.testing.assert_array_equal(dec.get_hard(), orc.hard())
            ran += 1
    assert ran >= len(CASES) * 3


POOL_PATHS = {
    "words_generic": dict(pairs=0, uniform=0),
    "words_generic_p1": dict(pairs=0, uniform=0, words_per_block=1),  # one word per workgroup
    "words_uniform_g1": dict(pairs=0, uniform=1, batch=1),
    "words_uniform_g2": dict(pairs=0, uniform=1, batch=2),
    "words_uniform_g4": dict(pairs=0, uniform=1, batch=4),
    "pairs": dict(pairs=1, words_per_block=2, threads=256),
    "pairs_p1": dict(pairs=1, words_per_block=1, threads=256),  # the one-pair-per-workgroup instantiations
}


@pytest.mark.parametrize("stop_rule", [L.STOP_SYNDROME, L.STOP_NONE], ids=["syndrome", "none"])
@pytest.mark.parametrize("path", list(POOL_PATHS))
def test_frame_queue_at_boundary_values(path, stop_rule, gpu_available):
    tuning = POOL_PATHS[path]
    ran = 0
    for bc in codes().values():
        single = len(set(bc.deg.tolist())) == 1
        if tuning.get("uniform") == 1 and not (single and 6 <= bc.deg[0] <= 8):
            continue
        for variant, table in CASES:
            if variant == L.OMS and not single:
                continue  # the frame queue takes OMS on one degree group only
            say(f"k_pool {path} stop={stop_rule}: code {bc.name} N={bc.n} F={bc.F} {case_name(variant, table)}")
            pool = L.PoolDecoder(bc.code, variant, ITERS, table=table, stop_rule=stop_rule, **tuning)
            cfg = pool.config()
            assert cfg["pairs"] == tuning["pairs"]
            if tuning.get("uniform") == 1:
                assert cfg["uniform_deg"] == bc.deg[0] and cfg["batch"] == tuning["batch"]
            hard, its = pool.decode(bc.llr)
            if stop_rule == L.STOP_NONE:
                # every frame runs ITERS iterations: the word decoder's result
                orc = O.OracleDecoder(bc.deg, bc.pos, bc.n, bc.F // 4, variant, q1_split=bc.code.q1_split,
                                      early_term=0, table=table)
                orc.load(bc.llr)
                orc.iterate(ITERS, threads=THREADS)
                np.testing.assert_array_equal(its, np.full(bc.F, ITERS, np.int32))
                np.testing.assert_array_equal(hard.reshape(-1), orc.hard(), err_msg=bc.name)
            else:
                # OMS on one degree group has no first-iteration exception: it is OMS_v2
                ov = L.OMS_V2 if variant == L.OMS else variant
                ref_h, ref_i = O.decode_frames(bc.deg, bc.pos, bc.n, ov, bc.llr, ITERS, table=table,
                                               threads=THREADS)
                np.testing.assert_array_equal(its, ref_i, err_msg=f"{bc.name}: iterations")
                np.testing.assert_array_equal(hard, ref_h, err_msg=f"{bc.name}: decisions")
            ran += 1
    assert ran >= len(CASES) * 3
