"""Learned tables on the GPU against the oracle.

* the shadow decoder's forward pass is bit-identical to the oracle's (same
  float operations, no contraction);
* loss and gradient agree with the oracle (reverse mode; summation order
  differs, so to float tolerance) -- and the oracle's gradient is itself
  pinned by finite differences (tests/test_shadow_oracle.py);
* the per-frame-Eb/N0 training channel equals the oracle's;
* Adam training lowers the held-out loss, and the quantised table is used
  bit-exactly by the integer decoder (LOMS against its oracle model);
* the committed trained table beats plain OMS on identical frames (what the
  learned part gains over the method without it)."""
import numpy as np
import pytest

import fast_ldpc_decoder_for_gpu_fixed_amd as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def code36():
    return L.code_3_6(4000, seed=1)


@pytest.mark.parametrize("mode", [0, 1])
def test_shadow_forward_and_gradient_match_oracle(mode, gpu_available):
    code = code36()
    deg, pos = code.export()
    F, T = 16, 4
    llr = O.channel_frames(4000, 0, F, 11, 0, O.sigma(1.75, 0.5), 0)
    rng = np.random.default_rng(mode)
    params = (rng.uniform(0.4, 2.2, (T, 7)) if mode == 0 else rng.uniform(0.5, 0.95, (T, 7))).astype(np.float32)
    tr = L.Trainer(code, mode, T, 7, F)
    d_llr = L.DeviceBuffer(llr.nbytes)
    d_llr.upload(llr)
    d_post = L.DeviceBuffer(F * 4000 * 4)
    loss, grad = tr.loss_grad(params, d_llr.ptr, True, d_post.ptr)
    L.synchronize()
    post = d_post.download(np.empty(F * 4000, np.float32)).reshape(F, 4000)
    oloss, opost = O.shadow_forward(deg, pos, 4000, mode, params, llr)
    np.testing.assert_array_equal(post, opost)
    oloss2, ograd = O.shadow_loss_grad(deg, pos, 4000, mode, params, llr)
    assert loss == pytest.approx(oloss, rel=1e-4)
    np.testing.assert_allclose(grad, ograd, rtol=2e-3, atol=1e-7)
    assert np.abs(grad[:, 6]).min() > 0


def test_range_channel_matches_oracle(gpu_available):
    N, F = 4000, 32
    d = L.DeviceBuffer(N * F)
    L.channel_generate_range(d.ptr, N, F, 5, 1.5, 2.5, 0.5, first_frame=1000)
    L.synchronize()
    got = d.download(np.empty(N * F, np.int8))
    np.testing.assert_array_equal(got, O.channel_frames_range(N, 1000, F, 5, 0, 1.5, 2.5, 0.5, 0))


def test_training_lowers_loss_and_table_is_bit_exact(gpu_available):
    code = code36()
    deg, pos = code.export()
    T, F = 6, 256
    tr = L.Trainer(code, L.LEARN_OFFSET, T, 7, F)
    init = np.ones((T, 7), np.float32)
    params, hist = tr.fit(init, steps=30, lr=5e-2, ebn0=(1.5, 2.5), seed=3, first_frame=0)
    assert np.isfinite(hist).all()
    # held-out frames (a different part of the stream)
    d_llr = L.DeviceBuffer(F * 4000)
    L.channel_generate_range(d_llr.ptr, 4000, F, 3, 1.5, 2.5, 0.5, first_frame=10_000_000)
    before, _ = tr.loss_grad(init, d_llr.ptr, False)
    after, _ = tr.loss_grad(params, d_llr.ptr, False)
    assert after < before, (before, after, params[:, 6])
    table = L.quantize_offsets(params)
    # the integer decoder with the quantised table == its oracle model
    W = 16
    llr = O.channel_frames(4000, 0, 4 * W, 21, 0, O.sigma(2.0, 0.5), 0)
    dec = L.make_decoder(code, W, L.LOMS, early_term=0)
    dec.set_table(table)
    out = dec.decode_bytes(llr, T)
    orc = O.OracleDecoder(deg, pos, 4000, W, O.LOMS, early_term=0, table=table)
    orc.load(llr)
    orc.iterate(T, threads=8)
    np.testing.assert_array_equal(out, orc.hard())


def test_simulate_matches_direct_decode(gpu_available):
    code = code36()
    W = 256
    dec = L.make_decoder(code, W, L.OMS, early_term=0)
    r = dec.simulate(1.5, 3 * 4 * W, 10, seed=8, snr_idx=2)
    assert r["frames"] == 3 * 4 * W and r["avg_iterations"] == 10
    # redo batch by batch by hand
    sig = L.channel_sigma(1.5, code.rate)
    d_llr, d_hard, cnt = L.DeviceBuffer(4 * W * 4000), L.DeviceBuffer(4 * W * 4000), L.DeviceBuffer(24)
    fe = be = 0
    for k in range(3):
        L.channel_generate(d_llr.ptr, 4000, 4 * W, 8, 2, sig, first_frame=k * 4 * W)
        dec.decode_device(d_llr.ptr, d_hard.ptr, 10)
        L.synchronize()
        f, b, _ = O.count_errors(d_hard.download(np.empty(4 * W * 4000, np.uint8)), 4000, code.n_info)
        fe, be = fe + f, be + b
    assert (r["frame_errors"], r["bit_errors"]) == (fe, be)
    et = L.make_decoder(code, W, L.OMS, early_term=1)
    r2 = et.simulate(2.5, 4 * W, 20, seed=8, snr_idx=2)
    assert 1 < r2["avg_iterations"] < 20


def test_committed_learned_table_beats_plain_oms(gpu_available):
    """What the trained table gains: on identical device-channel frames at
    1.75 dB, the learned-offset decoder has far fewer frame errors than plain
    OMS, both at 10 fixed iterations and in the frame-queue decoder (max 20,
    syndrome stop), and it stops earlier on average."""
    from fast_ldpc_decoder_for_gpu_fixed_amd.learned import load_table
    from pathlib import Path
    variant, table = load_table(str(Path(L.__file__).parent / "tables" / "loms_3_6_n4000.json"))
    assert variant == L.LOMS
    code = code36()
    F = 1 << 16
    fixed = {}
    for name, v in (("plain", L.OMS), ("learned", L.LOMS)):
        dec = L.make_decoder(code, F // 4, v, early_term=0)
        if v == L.LOMS:
            dec.set_table(table)
        fixed[name] = dec.simulate(1.75, F, 10, seed=31, snr_idx=7)
    assert fixed["learned"]["frame_errors"] < 0.6 * fixed["plain"]["frame_errors"], fixed
    pool = {}
    for name, v in (("plain", L.OMS), ("learned", L.LOMS)):
        p = L.PoolDecoder(code, v, 20, table=table if v == L.LOMS else None)
        pool[name] = p.simulate(1.75, F, seed=31, snr_idx=7)
    assert pool["learned"]["frame_errors"] < 0.5 * pool["plain"]["frame_errors"], pool
    assert pool["learned"]["avg_iterations"] < pool["plain"]["avg_iterations"]
