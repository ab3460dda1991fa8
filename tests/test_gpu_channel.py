"""Device channel and device error counters against the oracle.

Channel: the same (seed, Eb/N0 index, sample) gives the same quantised LLR
on the GPU (Philox + double-precision Box-Muller) and in the oracle; both
quantizers, both layouts.  Counters: bit/frame errors over the first N-K
positions (src/ber_analyzer/CErrorAnalyzer.cpp:141-157) and the syndrome
check equal the oracle's host scan."""
import numpy as np
import pytest

import fast_ldpc_decoder_for_gpu_fixed_amd as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("quant", [0, 1])
@pytest.mark.parametrize("N", [4000, 1001])
def test_channel_matches_oracle(quant, N, gpu_available):
    frames, f0, seed, idx = 64, 12345, 77, 3
    sig = L.channel_sigma(2.0, 0.5)
    assert sig == pytest.approx(O.sigma(2.0, 0.5), rel=1e-15)
    d = L.DeviceBuffer(frames * N)
    L.channel_generate(d.ptr, N, frames, seed, idx, sig, quant=quant, first_frame=f0)
    L.synchronize()
    got = d.download(np.empty(frames * N, np.int8))
    np.testing.assert_array_equal(got, O.channel_frames(N, f0, frames, seed, idx, sig, quant))


def test_channel_word_layout(gpu_available):
    N, frames = 4000, 128
    sig = L.channel_sigma(1.5, 0.5)
    d = L.DeviceBuffer(frames * N)
    L.channel_generate(d.ptr, N, frames, 9, 1, sig, layout=1)
    L.synchronize()
    words = d.download(np.empty(N * frames // 4, np.uint32))
    np.testing.assert_array_equal(words, O.interleave(O.channel_frames(N, 0, frames, 9, 1, sig, 0), N))


def test_channel_statistics_full_batch(gpu_available):
    """2^16 frames: the fraction of positive LLRs equals Q(1/sigma)."""
    import math
    N, frames = 4000, 1 << 16
    sig = L.channel_sigma(2.0, 0.5)
    d = L.DeviceBuffer(frames * N)
    L.channel_generate(d.ptr, N, frames, 1, 0, sig)
    L.synchronize()
    x = d.download(np.empty(frames * N, np.int8))
    # round(16y/sigma^2) >= 1  <=>  y >= sigma^2/32;  >= 0  <=>  y > -sigma^2/32
    Q = lambda t: 0.5 * math.erfc(t / math.sqrt(2))
    for frac, p in ((float((x > 0).mean()), Q((1 + sig ** 2 / 32) / sig)),
                    (float((x >= 0).mean()), Q((1 - sig ** 2 / 32) / sig))):
        assert abs(frac - p) < 6 * math.sqrt(p * (1 - p) / x.size)


def test_device_counters_match_oracle(gpu_available):
    code = L.code_3_6(4000, seed=1)
    deg, pos = code.export()
    W = 64
    frames = 4 * W
    sig = L.channel_sigma(1.25, 0.5)
    d_llr = L.DeviceBuffer(frames * 4000)
    d_hard = L.DeviceBuffer(frames * 4000)
    L.channel_generate(d_llr.ptr, 4000, frames, 5, 0, sig)
    dec = L.make_decoder(code, W, L.OMS)
    dec.decode_device(d_llr.ptr, d_hard.ptr, 10)
    cnt = L.DeviceBuffer(3 * 8)
    per_bits = L.DeviceBuffer(frames * 4)
    per_ok = L.DeviceBuffer(frames)
    dec.count_errors(d_hard.ptr, frames, code.n_info, cnt.ptr, per_bits.ptr, per_ok.ptr, stream=dec.stream)
    L.synchronize()
    c = cnt.download(np.empty(3, np.uint64))
    hard = d_hard.download(np.empty(frames * 4000, np.uint8))
    fe, be, per = O.count_errors(hard, 4000, code.n_info)
    ok = O.syndrome_ok(hard, 4000, deg, pos)
    assert fe > 0 and (ok == 0).any()  # 1.25 dB: some frames fail
    assert (int(c[0]), int(c[1]), int(c[2])) == (be, fe, int((ok == 0).sum()))
    np.testing.assert_array_equal(per_bits.download(np.empty(frames, np.uint32)), per)
    np.testing.assert_array_equal(per_ok.download(np.empty(frames, np.uint8)), ok)
    # accumulate adds
    dec.count_errors(d_hard.ptr, frames, code.n_info, cnt.ptr, accumulate=True, stream=dec.stream)
    L.synchronize()
    c2 = cnt.download(np.empty(3, np.uint64))
    assert (int(c2[0]), int(c2[1])) == (2 * be, 2 * fe)
