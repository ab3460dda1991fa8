"""GPU parity of the persistent frame-queue decoder (ldpc_pool_*) against the
oracle's per-frame stop decode: decisions and iteration count of every frame
bit-exact, for every variant, with slots refilled many times over (few
workgroups, many frames), on codes that exercise each degree class, a frame
length that is not a multiple of four, learned tables and the default grid."""
import os
from pathlib import Path

import numpy as np
import pytest

import fast_ldpc_decoder_for_gpu_fixed_amd as L
from oracle import oracle as O

ROOT = Path(__file__).resolve().parents[1]

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def mixed_llr(n, frames, seed, lo=1.0, hi=3.5, rate=0.5):
    """All-zero codeword, BPSK bit0 -> -1, per-frame Eb/N0 in [lo, hi]."""
    rng = np.random.default_rng(seed)
    ebn0 = rng.uniform(lo, hi, size=(frames, 1))
    sigma = np.sqrt(1.0 / (2 * rate * 10 ** (ebn0 / 10)))
    y = -1.0 + sigma * rng.standard_normal((frames, n))
    return np.clip(np.round(16.0 * y / sigma ** 2), -127, 127).astype(np.int8)


def table_for(variant, iters=20, degs=33, seed=5):
    if variant not in (L.LOMS, L.LNMS):
        return None
    rng = np.random.default_rng(seed)
    lo, hi = (0, 3) if variant == L.LOMS else (5, 8)
    return rng.integers(lo, hi + 1, size=(iters, degs)).astype(np.int8)


def check_pool(code, variant, llr, max_it, workgroups=0, stop_rule=L.STOP_SYNDROME, table=None, **tuning):
    deg, pos = code.export()
    table = table_for(variant) if table is None else table
    pool = L.PoolDecoder(code, variant, max_it, table=table, stop_rule=stop_rule, **tuning)
    hard, its = pool.decode(llr, workgroups=workgroups)
    ref_h, ref_i = O.decode_frames(deg, pos, code.n_vars, L.OMS_V2 if variant == L.OMS else variant, llr, max_it,
                                   table=table, threads=THREADS, stop_rule=stop_rule)
    np.testing.assert_array_equal(its, ref_i, err_msg="iterations per frame")
    np.testing.assert_array_equal(hard, ref_h, err_msg="decisions")
    return its


_CODES = {}


def code(name):
    if name not in _CODES:
        _CODES[name] = {
            "c36_1000": lambda: L.code_3_6(1000, seed=3),
            "c36_1002": lambda: L.code_3_6(1002, seed=4),
            "qc_4x24_z24": lambda: L.Code.qc(4, 24, 24, 0.9, seed=2),
            "qc_12x24_z27": lambda: L.Code.qc(12, 24, 27, 0.9, seed=2),
            "qc_12x24_z27_d9": lambda: L.Code.qc(12, 24, 27, 0.35, seed=3),  # degrees 3..9
        }[name]()
    return _CODES[name]


@pytest.mark.parametrize("stop_rule", [L.STOP_SYNDROME, L.STOP_V2C_PARITY], ids=["syndrome", "v2c"])
@pytest.mark.parametrize("variant", [L.MS, L.OMS, L.OMS_V2, L.NMS, L.NMS2, L.LOMS, L.LNMS],
                         ids=lambda v: L.VARIANT_NAMES[v])
def test_pool_parity_refills(variant, stop_rule, gpu_available):
    c = code("c36_1000")
    llr = mixed_llr(c.n_vars, 2000, seed=variant)
    its = check_pool(c, variant, llr, 20, workgroups=3, stop_rule=stop_rule)
    assert (its < 20).any()


@pytest.mark.parametrize("name", ["c36_1002", "qc_4x24_z24", "qc_12x24_z27", "qc_12x24_z27_d9"])
@pytest.mark.parametrize("variant", [L.OMS_V2, L.NMS2, L.LOMS])
def test_pool_parity_codes(name, variant, gpu_available):
    c = code(name)
    rate = 1.0 - c.n_checks / c.n_vars
    llr = mixed_llr(c.n_vars, 1500, seed=7, lo=1.5, hi=4.0, rate=rate)
    check_pool(c, variant, llr, 25, workgroups=5)
    check_pool(c, variant, llr, 25, workgroups=5, stop_rule=L.STOP_V2C_PARITY)


@pytest.mark.parametrize("P,T,uniform,batch", [(1, 64, 1, 1), (2, 128, 1, 2), (4, 256, 1, 4), (16, 512, 1, 2),
                                              (2, 256, 0, 0), (16, 256, 0, 0),
                                              (2, 192, 1, 2), (4, 320, 1, 1), (2, 192, 0, 0)])
def test_pool_tunings_and_default_grid(P, T, uniform, batch, gpu_available):
    c = code("c36_1000")
    llr = mixed_llr(c.n_vars, 1200, seed=P)
    pool = L.PoolDecoder(c, L.OMS_V2, 15, words_per_block=P, threads=T, uniform=uniform, batch=batch)
    cfg = pool.config()
    assert cfg["uniform_deg"] == (6 if uniform else 0)
    check_pool(c, L.OMS_V2, llr, 15, workgroups=0, words_per_block=P, threads=T, uniform=uniform, batch=batch)


@pytest.mark.parametrize("name,P,T", [("qc_12x24_z27_d9", 2, 192), ("qc_4x24_z24", 1, 192), ("c36_1000", 2, 320)])
@pytest.mark.parametrize("fmt", ["words", "pairs"])
def test_pool_workgroups_in_whole_waves(name, P, T, fmt, gpu_available):
    """Workgroup sizes that are whole waves but not powers of two (the
    autotuner tries 192, 320, 384, 768): decisions and iteration counts on the
    generic word path and in pair mode."""
    c = code(name)
    rate = 1.0 - c.n_checks / c.n_vars
    llr = mixed_llr(c.n_vars, 1500, seed=T, lo=1.5, hi=4.0, rate=rate)
    tuning = {"pairs": 1 if fmt == "pairs" else 0, "words_per_block": P, "threads": T}
    assert L.PoolDecoder(c, L.OMS_V2, 25, **tuning).config()["threads"] == T
    check_pool(c, L.OMS_V2, llr, 25, workgroups=3, **tuning)


@pytest.mark.parametrize("variant", [L.NMS2, L.LOMS, L.LNMS], ids=lambda v: L.VARIANT_NAMES[v])
def test_pool_generic_path_learned(variant, gpu_available):
    """The same frames through the generic (any degree) path."""
    c = code("c36_1000")
    llr = mixed_llr(c.n_vars, 1000, seed=11)
    check_pool(c, variant, llr, 20, workgroups=2, uniform=0)


@pytest.mark.parametrize("fmt", ["words", "pairs"])
@pytest.mark.parametrize("stop_rule", [L.STOP_SYNDROME, L.STOP_V2C_PARITY], ids=["syndrome", "v2c"])
@pytest.mark.parametrize("name", ["qc_2x4_z3", "two_checks"])
def test_pool_tiny_codes(name, stop_rule, fmt, gpu_available):
    """Codes with fewer edges than a check gather reads (8 entries for degrees
    up to 8): those reads run past the last edge into the arrays' padding;
    decisions and iteration counts stay the oracle's, on the generic word path
    and in pair mode."""
    c = (L.Code.qc(2, 4, 3, 0.9, seed=1) if name == "qc_2x4_z3" else
         L.Code.from_table(6, np.array([3, 3], np.int32), np.array([0, 1, 2, 2, 3, 4], np.uint32)))
    assert c.n_edges <= 24
    llr = mixed_llr(c.n_vars, 600, seed=4, lo=0.0, hi=4.0)
    tuning = {"pairs": 1} if fmt == "pairs" else {"uniform": 0}
    check_pool(c, L.OMS_V2, llr, 12, workgroups=2, stop_rule=stop_rule, **tuning)


def test_pool_small_queue_and_reuse(gpu_available):
    """Fewer frames than slots, then a second call on the same pool."""
    c = code("c36_1000")
    pool = L.PoolDecoder(c, L.MS, 10)
    deg, pos = c.export()
    for F, seed in ((3, 1), (777, 2)):
        llr = mixed_llr(c.n_vars, F, seed=seed)
        hard, its = pool.decode(llr)
        rh, ri = O.decode_frames(deg, pos, c.n_vars, L.MS, llr, 10)
        np.testing.assert_array_equal(its, ri)
        np.testing.assert_array_equal(hard, rh)


def test_pool_errors(gpu_available):
    grouped = L.Code.qc(12, 24, 27, 0.9, seed=2)
    if grouped.q1_split < grouped.n_checks:
        with pytest.raises(L.LdpcError):
            L.PoolDecoder(grouped, L.OMS, 10)
    c = code("c36_1000")
    with pytest.raises(L.LdpcError):
        L.PoolDecoder(c, L.MS, 0)
    with pytest.raises(L.LdpcError):
        L.PoolDecoder(c, L.MS, 10, words_per_block=3)
    p = L.PoolDecoder(c, L.LOMS, 10)
    d = L.DeviceBuffer(c.n_vars * 4)
    with pytest.raises(L.LdpcError):  # learned variant without a table
        p.decode_device(d.ptr, 4, d.ptr)
    with pytest.raises(L.LdpcError):  # an empty call
        L.PoolDecoder(c, L.MS, 10).decode_device(d.ptr, 0, d.ptr)


def test_pool_simulate_matches_direct_decode(gpu_available):
    """Device channel -> pool -> device counters equals decoding the same
    channel frames by hand and counting them with the oracle; the frames are
    the ones the word decoder's simulate draws (same stream positions)."""
    c = L.code_3_6(4000, seed=1)
    F = 3000
    pool = L.PoolDecoder(c, L.OMS, 20)
    r = pool.simulate(1.75, F, seed=8, snr_idx=2, batch_frames=1024)  # three launches, the last partial
    sig = L.channel_sigma(1.75, c.rate)
    d_llr, d_hard, d_it = L.DeviceBuffer(F * 4000), L.DeviceBuffer(F * 4000), L.DeviceBuffer(4 * F)
    L.channel_generate(d_llr.ptr, 4000, F, 8, 2, sig)
    pool.decode_device(d_llr.ptr, F, d_hard.ptr, d_it.ptr)
    L.synchronize()
    hard = d_hard.download(np.empty(F * 4000, np.uint8))
    its = d_it.download(np.empty(F, np.int32))
    f, b, _ = O.count_errors(hard, 4000, c.n_info)
    assert (r["frame_errors"], r["bit_errors"]) == (f, b)
    assert r["avg_iterations"] == pytest.approx(its.mean(), rel=1e-12)
    # and these are the word decoder's frames: 20 fixed iterations vs per-frame stop
    llr = d_llr.download(np.empty((F, 4000), np.int8))
    ref_h, ref_i = O.decode_frames(*c.export(), 4000, L.OMS_V2, llr[:200], 20, threads=THREADS)
    np.testing.assert_array_equal(its[:200], ref_i)
    np.testing.assert_array_equal(hard.reshape(F, 4000)[:200], ref_h)


# ---------------------------------------------------------------------------
# pair mode: 2 frames per 16-bit posterior entry (long frames whose 4-frame
# word block does not fit LDS; forced here on short codes as well)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("stop_rule", [L.STOP_SYNDROME, L.STOP_V2C_PARITY], ids=["syndrome", "v2c"])
@pytest.mark.parametrize("variant", [L.MS, L.OMS_V2, L.NMS, L.NMS2, L.LOMS, L.LNMS],
                         ids=lambda v: L.VARIANT_NAMES[v])
def test_pool_pairs_parity(variant, stop_rule, gpu_available):
    c = code("c36_1000")
    llr = mixed_llr(c.n_vars, 1500, seed=40 + variant)
    pool = L.PoolDecoder(c, variant, 20, pairs=1, words_per_block=4, threads=256, table=table_for(variant),
                         stop_rule=stop_rule)
    assert pool.config()["pairs"] == 1 and pool.config()["uniform_deg"] == 0
    check_pool(c, variant, llr, 20, workgroups=3, stop_rule=stop_rule, pairs=1, words_per_block=4, threads=256)


@pytest.mark.parametrize("name", ["c36_1002", "qc_4x24_z24", "qc_12x24_z27"])
@pytest.mark.parametrize("P,T", [(1, 64), (8, 512)])
def test_pool_pairs_codes(name, P, T, gpu_available):
    c = code(name)
    if c.max_deg > 8 and T > 256:
        T = 256
    rate = 1.0 - c.n_checks / c.n_vars
    llr = mixed_llr(c.n_vars, 1200, seed=8, lo=1.5, hi=4.0, rate=rate)
    check_pool(c, L.OMS_V2, llr, 25, workgroups=0, pairs=1, words_per_block=P, threads=T)


def ira_llr(c, frames, ebn0, seed):
    """The reference-compatible quantizer (trunc(8y) clamped to +-31)."""
    from oracle import oracle as O
    return O.channel_frames(c.n_vars, 0, frames, seed, 0, O.sigma(ebn0, c.n_info / c.n_vars), 1).reshape(frames, -1)


def test_pool_pairs_ira_small(gpu_available):
    c = L.Code.ira(600, 240, 8, 3, 2).reorder_by_colour()
    llr = ira_llr(c, 800, 1.5, 3)
    for v in (L.OMS_V2, L.LOMS):
        check_pool(c, v, llr, 30, workgroups=4, pairs=1, words_per_block=2, threads=256)


def test_pool_ira_n64800_auto_pairs(gpu_available):
    """The long IRA code: a 4-frame word block (259 KB) does not fit LDS, so
    the decoder picks pair mode by itself."""
    c = L.Code.ira(32400, 12960, 8, 3, 1).reorder_by_colour()
    llr = ira_llr(c, 24, 1.3, 5)
    pool = L.PoolDecoder(c, L.OMS_V2, 50)
    cfg = pool.config()
    assert cfg["pairs"] == 1 and cfg["words_per_block"] == 1
    its = check_pool(c, L.OMS_V2, llr, 50)
    assert its.max() < 50


def test_pool_ira_n64800_learned_table(gpu_available):
    """The IRA code with the committed learned table (the bench's learned leg:
    one pair per workgroup, the learned-offset kernel) against the oracle."""
    from fast_ldpc_decoder_for_gpu_fixed_amd.learned import load_table
    variant, table = load_table(str(ROOT / "fast_ldpc_decoder_for_gpu_fixed_amd" / "tables" / "loms_ira_n64800.json"))
    c = L.Code.ira(32400, 12960, 8, 3, 1).reorder_by_colour()
    llr = ira_llr(c, 24, 1.1, 6)
    pool = L.PoolDecoder(c, variant, 50, table=table)
    assert pool.config()["pairs"] == 1 and pool.config()["words_per_block"] == 1
    check_pool(c, variant, llr, 50, table=table)


# ---------------------------------------------------------------------------
# size limits: the largest frames each posterior format holds in LDS
# (posterior block N*4*P bytes for words, N*2*P for pairs, at most 152 KiB)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n,pairs", [(38910, 0), (77820, 1)], ids=["words_N38910", "pairs_N77820"])
def test_pool_largest_frames(n, pairs, gpu_available):
    c = L.code_3_6(n, seed=9)
    pool = L.PoolDecoder(c, L.OMS_V2, 12, pairs=pairs, words_per_block=1)
    cfg = pool.config()
    assert cfg["pairs"] == pairs and cfg["words_per_block"] == 1
    llr = mixed_llr(n, 24, seed=3, lo=1.8, hi=3.0)
    check_pool(c, L.OMS_V2, llr, 12, pairs=pairs, words_per_block=1)


def test_pool_rejects_blocks_beyond_lds(gpu_available):
    c = L.code_3_6(38916, seed=9)  # N*4 just above the posterior budget
    with pytest.raises(L.LdpcError):
        L.PoolDecoder(c, L.OMS_V2, 10, pairs=0)
    assert L.PoolDecoder(c, L.OMS_V2, 10).config()["pairs"] == 1  # auto: pair mode


def test_pool_graph_replays_match_eager_and_oracle(gpu_available):
    """ldpc_pool_decode_graph: the captured step (queue-head reset kernel +
    frame-queue kernel) replayed several times, with eager work and
    synchronisations between replays, equals the eager decode and the oracle;
    new buffers, a new table or another frame count recapture."""
    c = L.code_3_6(4000, seed=1)
    deg, pos = c.export()
    F = 4096
    llr = mixed_llr(c.n_vars, F, seed=31, lo=1.5, hi=2.5)
    table = table_for(L.LOMS)
    pool = L.PoolDecoder(c, L.LOMS, 20, table=table)
    ref_h, ref_i = O.decode_frames(deg, pos, c.n_vars, L.LOMS, llr, 20, table=table, threads=THREADS)
    d_in, d_out, d_it = L.DeviceBuffer(llr.nbytes), L.DeviceBuffer(llr.nbytes), L.DeviceBuffer(4 * F)
    d_in.upload(llr)

    def run(graph, out=d_out, it=d_it, frames=F):
        out.upload(np.full(llr.nbytes, 0xAA, np.uint8))
        it.upload(np.full(F, -1, np.int32))
        pool.decode_device(d_in.ptr, frames, out.ptr, it.ptr, graph=graph)
        L.synchronize()
        return out.download(np.empty((F, c.n_vars), np.uint8)), it.download(np.empty(F, np.int32))

    for rep in range(3):
        h, i = run(graph=True)
        np.testing.assert_array_equal(i, ref_i, err_msg=f"graph replay {rep}")
        np.testing.assert_array_equal(h, ref_h, err_msg=f"graph replay {rep}")
        h, i = run(graph=False)
        np.testing.assert_array_equal(i, ref_i, err_msg=f"eager after replay {rep}")
        np.testing.assert_array_equal(h, ref_h)
    # other output buffers: recaptured
    o2, i2 = L.DeviceBuffer(llr.nbytes), L.DeviceBuffer(4 * F)
    h, i = run(graph=True, out=o2, it=i2)
    np.testing.assert_array_equal(i, ref_i)
    np.testing.assert_array_equal(h, ref_h)
    # fewer frames: recaptured, the rest of the buffer untouched
    h, i = run(graph=True, frames=F // 2)
    np.testing.assert_array_equal(i[:F // 2], ref_i[:F // 2])
    np.testing.assert_array_equal(h[:F // 2], ref_h[:F // 2])
    assert (i[F // 2:] == -1).all() and (h[F // 2:] == 0xAA).all()
    # a new table: recaptured with it
    t2 = table_for(L.LOMS, seed=6)
    pool.set_table(t2)
    r2_h, r2_i = O.decode_frames(deg, pos, c.n_vars, L.LOMS, llr, 20, table=t2, threads=THREADS)
    h, i = run(graph=True)
    np.testing.assert_array_equal(i, r2_i)
    np.testing.assert_array_equal(h, r2_h)


# ---------------------------------------------------------------------------
# queue order: least reliable frames first (ldpc_pool_set_order)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("fmt", ["words", "pairs"])
def test_pool_order_least_reliable_first_matches_oracle(fmt, gpu_available):
    """Frames queued noisiest first (a per-frame Eb/N0 spread, so the order
    differs from the input's): every frame's decisions and iteration count
    equal the oracle's and the input-order decode's, with few workgroups (many
    refills) and with the default grid."""
    c = code("qc_12x24_z27_d9")
    F = 1500
    llr = mixed_llr(c.n_vars, F, seed=41, lo=0.8, hi=3.2)
    tuning = dict(pairs=1 if fmt == "pairs" else 0, words_per_block=2)
    for wg in (3, 0):
        ref_i = check_pool(c, L.OMS_V2, llr, 30, workgroups=wg, **tuning)
        pool = L.PoolDecoder(c, L.OMS_V2, 30, order=L.ORDER_LEAST_RELIABLE_FIRST, **tuning)
        hard, its = pool.decode(llr, workgroups=wg)
        np.testing.assert_array_equal(its, ref_i, err_msg=f"ordered iterations, workgroups={wg}")
        deg, pos = c.export()
        ref_h, _ = O.decode_frames(deg, pos, c.n_vars, L.OMS_V2, llr, 30, threads=THREADS)
        np.testing.assert_array_equal(hard, ref_h, err_msg=f"ordered decisions, workgroups={wg}")
    assert len(set(ref_i.tolist())) > 3  # iteration counts spread: the order matters to the queue


def test_pool_order_ira_learned_and_graph(gpu_available):
    """The long IRA code in pair mode with its learned table, ordered, eager
    and as a replayed graph; a larger call regrows the order scratch (the
    graph is recaptured), a smaller one reuses it; switching the order back
    recaptures too.  Every result equals the oracle's."""
    from fast_ldpc_decoder_for_gpu_fixed_amd.learned import load_table
    variant, table = load_table(str(ROOT / "fast_ldpc_decoder_for_gpu_fixed_amd" / "tables" / "loms_ira_n64800.json"))
    c = L.Code.ira(32400, 12960, 8, 3, 1).reorder_by_colour()
    deg, pos = c.export()
    F = 40
    llr = ira_llr(c, F, 1.2, 8)
    ref_h, ref_i = O.decode_frames(deg, pos, c.n_vars, variant, llr, 50, table=table, threads=THREADS)
    pool = L.PoolDecoder(c, variant, 50, table=table, order=L.ORDER_LEAST_RELIABLE_FIRST)
    d_in, d_out, d_it = L.DeviceBuffer(llr.nbytes), L.DeviceBuffer(llr.nbytes), L.DeviceBuffer(4 * F)
    d_in.upload(llr)

    def run(frames, graph):
        d_out.upload(np.full(llr.nbytes, 0xAA, np.uint8))
        d_it.upload(np.full(F, -1, np.int32))
        pool.decode_device(d_in.ptr, frames, d_out.ptr, d_it.ptr, graph=graph)
        L.synchronize()
        h, i = d_out.download(np.empty((F, c.n_vars), np.uint8)), d_it.download(np.empty(F, np.int32))
        np.testing.assert_array_equal(i[:frames], ref_i[:frames], err_msg=f"frames={frames} graph={graph}")
        np.testing.assert_array_equal(h[:frames], ref_h[:frames], err_msg=f"frames={frames} graph={graph}")
        assert (i[frames:] == -1).all()

    for frames, graph in ((16, False), (16, True), (16, True), (F, True), (F, False), (24, True), (F, True)):
        run(frames, graph)
    pool.set_order(L.ORDER_INPUT)
    run(F, True)
    pool.set_order(L.ORDER_LEAST_RELIABLE_FIRST)
    run(F, True)
    with pytest.raises(L.LdpcError):
        pool.set_order(2)
