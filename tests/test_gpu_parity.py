"""GPU parity: the HIP decoder against the CPU oracle, through the C-ABI.

Bar: bit-exact posterior and message arrays after EVERY iteration, every
variant, plus hard decisions.  Sizes are chosen so the oracle finishes in
seconds; the full-size batch (2^16 frames) is compared whole: every decision
at 10 and 20 iterations for the five variants, and the complete posterior and
message arrays for OMS."""
import os

import numpy as np
import pytest

import fast_ldpc_decoder_for_gpu_fixed_amd as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def channel_llr(n, frames, ebn0=2.0, rate=0.5, seed=0, quant=0):
    """Host test input: all-zero codeword, BPSK bit0 -> -1, AWGN."""
    rng = np.random.default_rng(seed)
    sigma = np.sqrt(1.0 / (2 * rate * 10 ** (ebn0 / 10)))
    y = -1.0 + sigma * rng.standard_normal(frames * n, dtype=np.float32)
    if quant == 0:
        q = np.clip(np.round(16.0 * y / sigma ** 2), -127, 127)
    else:
        q = np.trunc(np.clip(8.0 * y, -31, 31))
    return q.astype(np.int8)


_CODES = {}


def code36():
    if "c36" not in _CODES:
        _CODES["c36"] = L.code_3_6(4000, seed=1)
    return _CODES["c36"]


def learned_table(variant, iters=20, degs=33, seed=5):
    rng = np.random.default_rng(seed)
    hi = 4 if variant == L.LOMS else 8
    lo = 0 if variant == L.LOMS else 4
    return rng.integers(lo, hi + 1, size=(iters, degs)).astype(np.int8)


def run_lockstep(code, W, variant, llr, iters, early_term=0, check_every=1, **tuning):
    deg, pos = code.export()
    dec = L.make_decoder(code, W, variant, early_term=early_term, **tuning)
    table = None
    if variant in (L.LOMS, L.LNMS):
        table = learned_table(variant)
        dec.set_table(table)
    orc = O.OracleDecoder(deg, pos, code.n_vars, W, variant, q1_split=code.q1_split,
                          early_term=early_term, table=table)
    orc.load(llr)
    dec.load(O.interleave(llr, code.n_vars).view(np.int8) if variant == L.OMS_V2 else llr)
    done = 0
    while done < iters:
        step = min(check_every, iters - done)
        dec.iterate(step)
        orc.iterate(step, threads=THREADS)
        done += step
        post, msg = dec.get_state()
        np.testing.assert_array_equal(post, orc.post, err_msg=f"posterior after iteration {done}")
        np.testing.assert_array_equal(msg, orc.msg, err_msg=f"messages after iteration {done}")
        if early_term:
            st, it = dec.stop_info()
            np.testing.assert_array_equal(st, orc.stopped)
            np.testing.assert_array_equal(it, orc.iters_run)
    np.testing.assert_array_equal(dec.get_hard(), orc.hard())
    return dec, orc


ALL = [L.MS, L.OMS, L.OMS_V2, L.NMS, L.NMS2, L.LOMS, L.LNMS]


@pytest.mark.parametrize("variant", ALL, ids=lambda v: L.VARIANT_NAMES[v])
def test_per_iteration_parity_3_6(variant, gpu_available):
    """(3,6) N=4000 code, 20 iterations, state compared after each one
    (covers the 10-iteration decode as a prefix)."""
    code = code36()
    W = 32
    llr = channel_llr(4000, 4 * W, ebn0=2.0, seed=variant)
    run_lockstep(code, W, variant, llr, 20)


@pytest.mark.parametrize("variant", [L.OMS, L.NMS2], ids=lambda v: L.VARIANT_NAMES[v])
def test_reference_quantizer_input(variant, gpu_available):
    code = code36()
    W = 16
    llr = channel_llr(4000, 4 * W, ebn0=1.5, seed=11, quant=1)
    run_lockstep(code, W, variant, llr, 10)


@pytest.mark.parametrize("variant", ALL, ids=lambda v: L.VARIANT_NAMES[v])
def test_one_shot_decode_10_and_20(variant, gpu_available):
    code = code36()
    deg, pos = code.export()
    W = 64
    llr = channel_llr(4000, 4 * W, ebn0=2.0, seed=100 + variant)
    for iters in (10, 20):
        dec = L.make_decoder(code, W, variant, early_term=0)
        table = None
        if variant in (L.LOMS, L.LNMS):
            table = learned_table(variant)
            dec.set_table(table)
        x = O.interleave(llr, 4000).view(np.int8) if variant == L.OMS_V2 else llr
        got = dec.decode_bytes(x, iters)
        orc = O.OracleDecoder(deg, pos, 4000, W, variant, q1_split=code.q1_split, early_term=0, table=table)
        orc.load(llr)
        orc.iterate(iters, threads=THREADS)
        np.testing.assert_array_equal(got, orc.hard())


def test_2nms_early_term_lockstep(gpu_available):
    """2NMS with its default per-word early stop (the reference compiles
    EARLY_TERM=1, src/decoder_2nms/cuda/CUDA_2NMS_SIMD.cu:17)."""
    code = code36()
    W = 32
    # reference quantizer at 2.5 dB: some words stop at iteration 6-7, some never
    llr = channel_llr(4000, 4 * W, ebn0=2.5, seed=9, quant=1)
    dec, orc = run_lockstep(code, W, L.NMS2, llr, 20, early_term=1)
    assert orc.stopped.any() and not orc.stopped.all()


def test_oms_early_term_lockstep(gpu_available):
    code = code36()
    W = 32
    llr = channel_llr(4000, 4 * W, ebn0=2.5, seed=19)
    _, orc = run_lockstep(code, W, L.OMS, llr, 15, early_term=1, check_every=3)
    assert orc.stopped.any()


def test_unordered_table_many_levels(gpu_available):
    """A table in plain generation order (28 dependency levels)."""
    code = L.code_3_6(4000, seed=1, colour=False)
    assert code.n_levels > 20
    W = 16
    llr = channel_llr(4000, 4 * W, seed=4)
    run_lockstep(code, W, L.OMS, llr, 6)


def test_mixed_degree_qc_code_q1(gpu_available):
    """Two+ check degrees: exercises the OMS first-iteration quirk on the
    checks after the leading degree group."""
    code = L.Code.qc(12, 24, 81, 0.35, seed=3)
    assert code.q1_split < code.n_checks
    W = 16
    for variant in (L.OMS, L.MS, L.NMS2):
        llr = channel_llr(code.n_vars, 4 * W, ebn0=1.5, seed=variant)
        run_lockstep(code, W, variant, llr, 8, check_every=2)


def random_code(n, c, dmin, dmax, seed):
    rng = np.random.default_rng(seed)
    deg = rng.integers(dmin, dmax + 1, size=c).astype(np.int32)
    pos = np.concatenate([rng.choice(n, size=d, replace=False) for d in deg]).astype(np.uint32)
    return L.Code.from_table(n, deg, pos)


@pytest.mark.parametrize("dmin,dmax", [(1, 8), (9, 16), (17, 32)])
def test_degree_classes(dmin, dmax, gpu_available):
    code = random_code(1000, 300, dmin, dmax, seed=dmax)
    W = 8
    for variant in (L.OMS, L.NMS2, L.LNMS):
        llr = channel_llr(1000, 4 * W, ebn0=2.0, seed=dmax + variant)
        run_lockstep(code, W, variant, llr, 4, threads=256)


def test_odd_length_frames_byte_path(gpu_available):
    code = random_code(1001, 500, 3, 7, seed=42)
    W = 4
    llr = channel_llr(1001, 4 * W, seed=1)
    run_lockstep(code, W, L.OMS, llr, 5)


@pytest.mark.parametrize("P,T,uniform", [(1, 64, 1), (2, 128, 1), (8, 256, 0), (4, 256, 1), (16, 1024, 0), (16, 512, 1),
                                         (4, 512, 0), (4, 256, 0)])
def test_tunings(P, T, uniform, gpu_available):
    code = code36() if P <= 8 else L.code_3_6(1000, seed=7)
    W = 32
    llr = channel_llr(code.n_vars, 4 * W, seed=P * T)
    dec, _ = run_lockstep(code, W, L.OMS, llr, 5, words_per_block=P, threads=T, uniform=uniform)
    assert dec.config()["uniform_deg"] == (6 if uniform else 0)


@pytest.mark.parametrize("batch", [1, 2, 4])
@pytest.mark.parametrize("variant", [L.OMS, L.NMS2, L.LOMS], ids=lambda v: L.VARIANT_NAMES[v])
def test_single_degree_kernel_batches(batch, variant, gpu_available):
    code = code36()
    W = 32
    llr = channel_llr(4000, 4 * W, seed=batch + 10 * variant)
    dec, _ = run_lockstep(code, W, variant, llr, 6, uniform=1, batch=batch, early_term=1 if variant == L.NMS2 else 0)
    assert dec.config()["uniform_deg"] == 6 and dec.config()["batch"] == batch


@pytest.mark.parametrize("d", [7, 8])
def test_single_degree_kernel_other_degrees(d, gpu_available):
    n = {7: 1400, 8: 1600}[d]
    code = L.Code.regular(n, 4, d, seed=d).reorder_by_colour() if d == 8 else L.Code.regular(n, 3, d, seed=d).reorder_by_colour()
    W = 16
    for batch in (1, 4):
        llr = channel_llr(n, 4 * W, seed=d * batch)
        dec, _ = run_lockstep(code, W, L.OMS, llr, 4, uniform=1, batch=batch)
        assert dec.config()["uniform_deg"] == d


def test_global_posterior_path(gpu_available):
    code = code36()
    W = 16
    llr = channel_llr(4000, 4 * W, seed=77)
    for variant in (L.OMS, L.NMS2):
        run_lockstep(code, W, variant, llr, 5, post_in_lds=0, words_per_block=4)
    # degrees 10..16: the generic kernel of that class, posterior in global memory
    qc = L.Code.qc(4, 24, 96, 0.35, seed=2)
    assert 8 < qc.max_deg <= 16
    llr = channel_llr(qc.n_vars, 4 * W, ebn0=3.5, rate=1.0 - qc.n_checks / qc.n_vars, seed=78)
    for variant in (L.OMS_V2, L.NMS2):
        run_lockstep(qc, W, variant, llr, 5, post_in_lds=0, words_per_block=4)


def full_batch_llr():
    """The drop-in's full-size batch: 2^16 frames (W = 16384) of the (3,6)
    N=4000 code at 2.0 dB."""
    if "full" not in _CODES:
        _CODES["full"] = channel_llr(4000, 4 * 16384, ebn0=2.0, seed=2024)
    return _CODES["full"]


@pytest.mark.parametrize("variant", [L.OMS, L.MS, L.NMS, L.NMS2, L.OMS_V2], ids=lambda v: L.VARIANT_NAMES[v])
def test_full_size_batch(variant, gpu_available):
    """2^16 frames at 2.0 dB, one decode() of the whole batch at 10 and at 20
    fixed iterations, no early stop (src/decoder_oms/CGPU_Decoder_OMS_SIMD.cu:97-149):
    every decision of every frame equals the oracle's."""
    code = code36()
    deg, pos = code.export()
    W, N = 16384, 4000
    llr = full_batch_llr()
    x = O.interleave(llr, N).view(np.int8) if variant == L.OMS_V2 else llr
    orc = O.OracleDecoder(deg, pos, N, W, variant, q1_split=code.q1_split, early_term=0)
    orc.load(llr)
    done = 0
    for iters in (10, 20):
        print(f"full batch {L.VARIANT_NAMES[variant]} W={W} iterations={iters}", flush=True)
        dec = L.make_decoder(code, W, variant, early_term=0)
        out = dec.decode_bytes(x, iters)
        del dec
        orc.iterate(iters - done, threads=THREADS)
        done = iters
        ref = orc.hard()
        bad = np.nonzero((out != ref).reshape(4 * W, N).any(axis=1))[0]
        assert bad.size == 0, f"{bad.size} frames differ after {iters} iterations, first {bad[:8].tolist()}"


def test_full_size_batch_state(gpu_available):
    """OMS, 2^16 frames, 20 iterations through the state API: the whole
    posterior [v][W] and message [e][W] arrays equal the oracle's
    (the reference's device_V / d_MSG_C_2_V, src/decoder_template/CGPUDecoder.cpp:14-38)."""
    code = code36()
    deg, pos = code.export()
    W, N = 16384, 4000
    llr = full_batch_llr()
    dec = L.make_decoder(code, W, L.OMS, early_term=0)
    dec.load(llr)
    dec.iterate(20)
    post, msg = dec.get_state()
    hard = dec.get_hard()
    del dec
    orc = O.OracleDecoder(deg, pos, N, W, L.OMS, q1_split=code.q1_split, early_term=0)
    orc.load(llr)
    orc.iterate(20, threads=THREADS)
    np.testing.assert_array_equal(post, orc.post, err_msg="posterior after 20 iterations")
    np.testing.assert_array_equal(msg, orc.msg, err_msg="messages after 20 iterations")
    np.testing.assert_array_equal(hard, orc.hard())


@pytest.mark.parametrize("variant,early", [(L.OMS, 0), (L.OMS_V2, 0), (L.NMS2, 1), (L.LOMS, 0)],
                         ids=["OMS", "OMS_v2", "2NMS_early", "LOMS"])
def test_host_decode_in_chunks_equals_device_path(variant, early, gpu_available):
    """2^16 frames through decode() from host buffers, which copies and
    decodes them in overlapped chunks of word blocks (OMS_v2: column ranges
    of its word-layout input), against the one-launch device path."""
    code = code36()
    W, N = 16384, 4000
    rng = np.random.default_rng(variant)
    x = rng.integers(-40, 41, size=4 * W * N, dtype=np.int8)
    x[: 4 * N * 64] = channel_llr(N, 4 * 64, seed=3).reshape(-1)  # some decodable frames as well
    dec = L.make_decoder(code, W, variant, early_term=early)
    assert dec.config()["words_per_block"] * 8 * 512 <= W  # the chunked path
    if variant == L.LOMS:
        with pytest.raises(L.LdpcError):  # no table yet: fails after the first copy was queued
            dec.decode_bytes(x, 12)
        dec.set_table(learned_table(variant))
    host = dec.decode_bytes(x, 12)
    d_in, d_out = L.DeviceBuffer(x.nbytes), L.DeviceBuffer(x.nbytes)
    d_in.upload(x)
    dec.decode_device(d_in.ptr, d_out.ptr, 12)
    L.synchronize()
    np.testing.assert_array_equal(host, d_out.download(np.empty(x.nbytes, np.uint8)))


def test_device_interleave_roundtrip(gpu_available):
    N, W = 4000, 96
    llr = channel_llr(N, 4 * W, seed=5)
    d_in = L.DeviceBuffer(llr.nbytes)
    d_w = L.DeviceBuffer(llr.nbytes)
    d_o = L.DeviceBuffer(llr.nbytes)
    d_in.upload(llr)
    lib = L.load()
    L.check(lib.ldpc_interleave_device(L.C.c_void_p(d_in.ptr), L.C.c_void_p(d_w.ptr), N, W, None), "interleave")
    L.check(lib.ldpc_deinterleave_hd_device(L.C.c_void_p(d_w.ptr), L.C.c_void_p(d_o.ptr), N, W, None), "deint")
    L.synchronize()
    words = d_w.download(np.empty(N * W, np.uint32))
    np.testing.assert_array_equal(words, O.interleave(llr, N))
    hd = d_o.download(np.empty(llr.size, np.uint8))
    np.testing.assert_array_equal(hd, (llr > 0).astype(np.uint8))


def test_error_paths(gpu_available):
    code = code36()
    dec = L.make_decoder(code, 8, L.OMS)
    llr = channel_llr(4000, 32)
    with pytest.raises(L.LdpcError):
        dec.decode_bytes(llr, 0)  # the reference would wrap to ~2^32 iterations
    with pytest.raises(L.LdpcError):
        dec.iterate(1)  # before load
    with pytest.raises(L.LdpcError):
        L.make_decoder(code, 6, L.OMS, words_per_block=4)  # W not a multiple of the block
    learned = L.make_decoder(code, 8, L.LOMS)
    with pytest.raises(L.LdpcError):
        learned.decode_bytes(llr, 5)  # no table yet
    with pytest.raises(L.LdpcError):
        learned.set_table(np.full((2, 7), -1, np.int8))
    # a table beyond the frame queue's limits is rejected and the decoder keeps its table
    learned.set_table(learned_table(L.LOMS))
    before = learned.decode_bytes(llr, 5)
    with pytest.raises(L.LdpcError):
        learned.set_table(np.zeros((65, 33), np.int8))
    with pytest.raises(L.LdpcError):
        learned.set_table(np.zeros((4, 34), np.int8))
    np.testing.assert_array_equal(learned.decode_bytes(llr, 5), before)
