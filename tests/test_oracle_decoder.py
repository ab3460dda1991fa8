"""Pins of the oracle's check-node update, schedule and layouts.

* check update: compared with a textbook statement written differently from
  the reference's min1/min2 bookkeeping -- for each edge, the minimum of the
  OTHER edges' |t| (capped at 127, the reference's initial minimum) and the
  parity of the OTHER edges' positive values -- brute-forced over small
  checks on values covering every saturation boundary;
* schedule: any reordering of the table that respects the dependency DAG
  leaves the state bit-identical (the GPU relies on this), and a reordering
  that breaks it changes the state (so the test has teeth);
* layouts: the closed form of the interleaver equals an emulation of the
  reference kernel's thread-index arithmetic (src/transpose/GPU_Transpose_uint8.cu);
* decoding sanity: noiseless frames decode to zero, message bounds hold.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O

EDGE_VALUES = [-128, -127, -100, -33, -32, -31, -30, -2, -1, 0, 1, 2, 30, 31, 32, 33, 96, 126, 127]


def s8(x):
    return ((int(x) + 128) % 256) - 128


def textbook(ts, variant, nocap=False, table_val=0):
    """Messages and posteriors of one check for one frame (ints)."""
    a = [abs(t) for t in ts]  # |-128| = 128
    mins = sorted(a)
    gmin = min(min(a), 127)
    out_m, out_p = [], []
    for j, t in enumerate(ts):
        others = a[:j] + a[j + 1:]
        m = min([127] + others)
        is_min = a[j] == gmin
        if variant == O.MS:
            mag = min(m, 31)
        elif variant in (O.OMS, O.OMS_V2):
            mag = max(m - 1, 0)
            if not nocap:
                mag = min(mag, 31)
        elif variant == O.NMS:
            mag = (3 * m) // 4
        elif variant == O.NMS2:
            mag = (7 * m) // 8 if is_min else (3 * m) // 4
        elif variant == O.LOMS:
            mag = min(max(m - table_val, 0), 31)
        elif variant == O.LNMS:
            mag = min((m * table_val) // 8, 31)
        pos_others = sum(1 for k, x in enumerate(ts) if k != j and x > 0)
        msg = mag if pos_others % 2 == 1 else -mag
        msg = s8(msg)
        out_m.append(msg)
        out_p.append(max(-128, min(127, t + msg)))
    del mins
    return out_m, out_p


def run_check(ts_lanes, variant, first=True, msgs_lanes=None, table=None, nocap_group=False):
    """One check of degree d on one word; ts_lanes: 4 lists of d posteriors."""
    d = len(ts_lanes[0])
    deg = [d, d] if nocap_group else [d]
    pos = list(range(d)) + ([d + j for j in range(d)] if nocap_group else [])
    n = d * len(deg)
    dec = O.OracleDecoder(deg, pos, n, 1, variant, q1_split=1 if nocap_group else -1,
                          early_term=0, table=table)
    post = np.zeros(n, np.uint32)
    for v in range(d):
        w = 0
        for k in range(4):
            w |= (ts_lanes[k][v] & 0xFF) << (8 * k)
        post[v] = w
        if nocap_group:
            post[d + v] = w
    dec.load_words(post)
    if not first:
        dec.first = False
        for v in range(d):
            w = 0
            for k in range(4):
                w |= (msgs_lanes[k][v] & 0xFF) << (8 * k)
            dec.msg[v] = w
            if nocap_group:
                dec.msg[d + v] = w
    dec.iterate(1)
    base = d if nocap_group else 0
    msgs = [[s8((int(dec.msg[base + v]) >> (8 * k)) & 0xFF) for v in range(d)] for k in range(4)]
    posts = [[s8((int(dec.post[base + v]) >> (8 * k)) & 0xFF) for v in range(d)] for k in range(4)]
    return msgs, posts


VARIANTS = [O.MS, O.OMS, O.OMS_V2, O.NMS, O.NMS2]


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("d", [2, 3])
def test_check_update_brute_force(variant, d):
    combos = list(itertools.product(EDGE_VALUES, repeat=d))
    # four frames per word: feed four consecutive combos per oracle call
    for i in range(0, len(combos) - 3, 4):
        lanes = [list(combos[i + k]) for k in range(4)]
        msgs, posts = run_check(lanes, variant)
        for k in range(4):
            em, ep = textbook(lanes[k], variant)
            assert msgs[k] == em and posts[k] == ep, (variant, lanes[k], msgs[k], em, posts[k], ep)


@pytest.mark.parametrize("variant", VARIANTS + [O.LOMS, O.LNMS])
@pytest.mark.parametrize("d", [4, 6, 8])
def test_check_update_random_with_messages(variant, d):
    rng = np.random.default_rng(100 + d + variant)
    table = None
    tv = 0
    if variant == O.LOMS:
        tv = 2
        table = np.full((1, 40), tv, np.int8)
    if variant == O.LNMS:
        tv = 5
        table = np.full((1, 40), tv, np.int8)
    for _ in range(150):
        post = [list(rng.choice(EDGE_VALUES, d)) if rng.random() < 0.5 else list(rng.integers(-128, 128, d))
                for _ in range(4)]
        msgs = [list(rng.integers(-40, 41, d)) for _ in range(4)]
        got_m, got_p = run_check(post, variant, first=False, msgs_lanes=msgs, table=table)
        for k in range(4):
            ts = [max(-128, min(127, p - m)) for p, m in zip(post[k], msgs[k])]
            em, ep = textbook(ts, variant, table_val=tv)
            assert got_m[k] == em and got_p[k] == ep


def test_oms_first_iteration_second_group_has_no_cap():
    """Quirk of src/decoder_oms/cuda/CUDA_OMS_SIMD.cu:115-116: first iteration,
    second degree group, the 31 clamp is skipped (MS and OMS_v2 clamp)."""
    lanes = [[100, -90, 80], [127, 127, 127], [-128, 50, 60], [5, 6, 7]]
    for variant, nocap in ((O.OMS, True), (O.OMS_V2, False), (O.MS, False)):
        msgs, posts = run_check(lanes, variant, nocap_group=True)
        for k in range(4):
            em, ep = textbook(lanes[k], variant, nocap=nocap)
            assert msgs[k] == em and posts[k] == ep
    # a clamp-free message above 31 really occurs
    msgs, _ = run_check(lanes, O.OMS, nocap_group=True)
    assert max(abs(m) for m in msgs[0]) > 31


# ---------------------------------------------------------------------------
# schedule invariance
# ---------------------------------------------------------------------------
def levels_py(deg, pos, n):
    last = [-1] * n
    out, e = [], 0
    for d in deg:
        vs = pos[e:e + d]
        e += d
        lv = max([last[v] + 1 for v in vs] + [0])
        for v in vs:
            last[v] = lv
        out.append(lv)
    return out


def reorder(deg, pos, order):
    off = np.concatenate([[0], np.cumsum(deg)])
    d2 = [int(deg[c]) for c in order]
    p2 = np.concatenate([pos[off[c]:off[c + 1]] for c in order]).astype(np.uint32)
    return np.array(d2, np.int32), p2


def msg_by_check(dec, deg):
    off = np.concatenate([[0], np.cumsum(deg)])
    W = dec.W
    return [dec.msg[off[c] * W:off[c + 1] * W].copy() for c in range(len(deg))]


@pytest.mark.parametrize("variant", [O.OMS, O.NMS2])
def test_dag_respecting_reorder_is_bit_identical(variant):
    import fast_ldpc_decoder_for_gpu_fixed_amd as L
    code = L.Code.regular(240, 3, 6, 5)
    deg, pos = code.export()
    lv = levels_py(list(deg), list(pos), 240)
    assert max(lv) + 1 == code.n_levels  # planner's level count matches this restatement
    # reverse the order inside every level: still respects every dependency
    order = sorted(range(len(deg)), key=lambda c: (lv[c], -c))
    deg2, pos2 = reorder(deg, pos, order)
    rng = np.random.default_rng(3)
    W = 8
    llr = rng.integers(-40, 30, size=4 * W * 240).astype(np.int8)
    a = O.OracleDecoder(deg, pos, 240, W, variant, early_term=0)
    b = O.OracleDecoder(deg2, pos2, 240, W, variant, early_term=0)
    a.load(llr)
    b.load(llr)
    a.iterate(6)
    b.iterate(6)
    np.testing.assert_array_equal(a.post, b.post)
    ma, mb = msg_by_check(a, deg), msg_by_check(b, deg2)
    for i, c in enumerate(order):
        np.testing.assert_array_equal(ma[c], mb[i])
    # negative control: plain reversal of the table breaks dependencies
    rev = list(range(len(deg)))[::-1]
    deg3, pos3 = reorder(deg, pos, rev)
    c3 = O.OracleDecoder(deg3, pos3, 240, W, variant, early_term=0)
    c3.load(llr)
    c3.iterate(6)
    assert not np.array_equal(a.post, c3.post)


def test_wave_balanced_levels_decode_like_the_table():
    """The decoders' wave-balanced schedule (checks with slack moved between
    levels) run in level order by the oracle gives the serial table's posterior
    and messages."""
    import fast_ldpc_decoder_for_gpu_fixed_amd as L
    code = L.Code.regular(240, 3, 6, 5)
    deg, pos = code.export()
    asap = np.array(levels_py(list(deg), list(pos), 240))
    lv = code.levels(8)
    assert lv.max() == asap.max() and not np.array_equal(lv, asap)  # some checks moved
    order = sorted(range(len(deg)), key=lambda c: (lv[c], c))
    deg2, pos2 = reorder(deg, pos, order)
    rng = np.random.default_rng(4)
    W = 8
    llr = rng.integers(-40, 30, size=4 * W * 240).astype(np.int8)
    a = O.OracleDecoder(deg, pos, 240, W, O.OMS, early_term=0)
    b = O.OracleDecoder(deg2, pos2, 240, W, O.OMS, early_term=0)
    a.load(llr)
    b.load(llr)
    a.iterate(5)
    b.iterate(5)
    np.testing.assert_array_equal(a.post, b.post)
    ma, mb = msg_by_check(a, deg), msg_by_check(b, deg2)
    for i, c in enumerate(order):
        np.testing.assert_array_equal(ma[c], mb[i])


# ---------------------------------------------------------------------------
# layout: emulate the reference interleaver's index arithmetic
# ---------------------------------------------------------------------------
def emulate_reference_interleaver(llr, N, W):
    """Thread-grid emulation of Interleaver_uint8 (GPU_Transpose_uint8.cu:80-130):
    grid (1, W/32), block (32, 32), int pointers over the byte buffers."""
    in32 = np.frombuffer(np.ascontiguousarray(llr, np.int8).tobytes() + b"\0" * 4096, np.uint32)
    out = np.zeros(N * W, np.uint32)
    n4 = N // 4
    for by in range(W // 32):
        for q in range(0, N, 128):
            s = np.zeros((4, 32, 32), np.uint32)
            for y in range(32):
                for t in range(32):
                    i = t + by * 32
                    bAddr = y * N + i % 32 + 32 * (i // 32) * N + q // 4
                    vals = [int(in32[bAddr + k * n4]) for k in range(4)]
                    for c in range(4):
                        s[c, t, y] = sum(((vals[k] >> (8 * c)) & 0xFF) << (8 * k) for k in range(4))
            for y in range(32):
                for t in range(32):
                    i = t + by * 32
                    base = 4 * W * y + i + q * W
                    if q // 4 + y < n4:
                        for c in range(4):
                            out[base + c * W] = s[c, y, t]
    return out


@pytest.mark.parametrize("N,W", [(256, 64), (136, 32)])
def test_interleave_closed_form_matches_reference_indexing(N, W):
    rng = np.random.default_rng(N)
    llr = rng.integers(-128, 128, size=4 * W * N).astype(np.int8)
    np.testing.assert_array_equal(O.interleave(llr, N), emulate_reference_interleaver(llr, N, W))
    hd = O.deinterleave_hd(O.interleave(llr, N), N)
    np.testing.assert_array_equal(hd, (llr > 0).astype(np.uint8))


# ---------------------------------------------------------------------------
# decoding sanity
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("variant", VARIANTS)
def test_noiseless_decodes_to_zero_and_bounds(variant):
    """MS/OMS clamp messages to 31 and decode near-noiseless frames to zero.
    NMS/2NMS have no clamp (src/decoder_nms/cuda/CUDA_NMS_SIMD.cu:73-85), so
    their messages reach 95/111, the int8 posterior saturates at -128 and
    post - msg loses magnitude every layer: the serial decoder oscillates on
    these frames.  Only their message bounds are checked."""
    import fast_ldpc_decoder_for_gpu_fixed_amd as L
    code = L.code_3_6(480, seed=2)
    deg, pos = code.export()
    W = 4
    rng = np.random.default_rng(1)
    llr = np.full(4 * W * 480, -31, np.int8)
    llr[rng.integers(0, llr.size, 200)] = 20  # a few flipped, still decodable
    dec = O.OracleDecoder(deg, pos, 480, W, variant, early_term=0)
    dec.load(llr)
    dec.iterate(10)
    m = dec.msg.view(np.int8).astype(int)
    bound = {O.MS: 31, O.OMS: 31, O.OMS_V2: 31, O.NMS: 95, O.NMS2: 111}[variant]
    assert np.abs(m).max() <= bound
    if variant in (O.MS, O.OMS, O.OMS_V2):
        assert not dec.hard().any()
        assert O.syndrome_ok(dec.hard(), 480, deg, pos).all()


def test_nms_saturation_trajectory_on_noiseless_frames():
    """Known answer for the unclamped NMS on an all-(-127) frame, derived by
    hand from src/decoder_nms/cuda/CUDA_NMS_SIMD.cu:40-132 and :160-220.
    Iteration 1 reads no messages; every |t| is 127 (or |-128| = 128 for a
    variable an earlier check already moved, above the 127 the minima start
    from), so every message is -trunc(0.75*127) = -95 and every posterior
    saturates at -128.  In iteration 2 the first check of the table sees
    t = -128 - (-95) = -33 on all its edges and sends -trunc(0.75*33) = -24:
    the magnitude lost to saturation is what makes the decoder oscillate."""
    import fast_ldpc_decoder_for_gpu_fixed_amd as L
    code = L.code_3_6(480, seed=2)
    deg, pos = code.export()
    W = 2
    dec = O.OracleDecoder(deg, pos, 480, W, O.NMS, early_term=0)
    dec.load(np.full(4 * W * 480, -127, np.int8))
    dec.iterate(1)
    assert (dec.post.view(np.int8) == -128).all()
    assert (dec.msg.view(np.int8) == -95).all()
    dec.iterate(1)
    first = dec.msg.view(np.int8).reshape(-1, 4 * W)[: int(deg[0])]  # [edge][frame] of check 0
    assert (first == -24).all()


def test_early_term_stops_on_zero_parity():
    import fast_ldpc_decoder_for_gpu_fixed_amd as L
    code = L.code_3_6(480, seed=2)
    deg, pos = code.export()
    W = 4
    llr = np.full(4 * W * 480, -31, np.int8)
    llr[4 * 480 * 2 + 5] = 31  # word 0 frame 2 has one wrong bit
    dec = O.OracleDecoder(deg, pos, 480, W, O.NMS2)
    dec.load(llr)
    dec.iterate(20)
    assert dec.stopped.all()
    assert (dec.iters_run == 2).all()  # earliest stop: end of the second iteration
    assert not dec.hard().any()
