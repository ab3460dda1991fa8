"""Pins of the oracle's channel: Philox4x32-10 known-answer vectors (Random123,
Salmon et al. SC'11), Gaussian moments, the uncoded BPSK error rate
Q(1/sigma), sigma from Eb/N0 (src/awgn_channel/CChanel_AWGN_SIMD.cu:63-73)
and both LLR quantizers (reference: trunc(clamp(8y, +-31)), :23-24)."""
import math

import numpy as np
import pytest

from oracle import oracle as O

KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    np.testing.assert_array_equal(O.philox(ctr, key), np.array(out, np.uint32))


def test_sigma_closed_form():
    for ebn0 in (1.0, 2.0, 3.0):
        for rate in (0.5, 5 / 6):
            assert O.sigma(ebn0, rate) == pytest.approx(math.sqrt(1 / (2 * rate * 10 ** (ebn0 / 10))), rel=1e-12)


def test_gauss_moments_and_uncoded_ber():
    n = 400
    frames = 500
    sig = O.sigma(2.0, 0.5)
    llr = O.channel_frames(n, 0, frames, 1234, 0, sig, 0).astype(np.int64)
    # uncoded BER of the channel: P(y > 0) = Q(1/sigma); sign of the LLR decides
    p = 0.5 * math.erfc((1 / sig) / math.sqrt(2))
    ber = float((llr > 0).mean())
    tol = 5 * math.sqrt(p * (1 - p) / llr.size)
    assert abs(ber - p) < tol
    g = np.array([O.lib().orc_gauss(7, 1, s) for s in range(20000)])
    assert abs(g.mean()) < 0.03 and abs(g.var() - 1) < 0.04


def test_quantizers():
    sig = O.sigma(1.0, 0.5)
    a = O.channel_frames(200, 3, 50, 99, 2, sig, 0).astype(int)
    b = O.channel_frames(200, 3, 50, 99, 2, sig, 1).astype(int)
    assert a.min() >= -127 and a.max() <= 127
    assert b.min() >= -31 and b.max() <= 31
    hi = O.channel_frames(200, 0, 20, 99, 2, O.sigma(9.0, 0.5), 0).astype(int)  # 16/sigma^2 ~ 127
    assert (hi == -127).mean() > 0.2 and hi.min() == -127
    # same noise sample drives both: signs agree wherever both are non-zero
    nz = (a != 0) & (b != 0)
    assert (np.sign(a[nz]) == np.sign(b[nz])).all()
    # position independence: frames generated alone equal the same frames in a batch
    c = O.channel_frames(200, 5, 2, 99, 2, sig, 0)
    np.testing.assert_array_equal(c, O.channel_frames(200, 3, 50, 99, 2, sig, 0)[2 * 200:4 * 200])


def test_count_errors_and_syndrome():
    hard = np.zeros((3, 10), np.uint8)
    hard[1, 2] = 1
    hard[2, 8] = 1  # outside the first n_info = 6 positions
    fe, be, per = O.count_errors(hard.ravel(), 10, 6)
    assert (fe, be) == (1, 1) and list(per) == [0, 1, 0]
    deg = np.array([2, 2], np.int32)
    pos = np.array([0, 1, 2, 8], np.uint32)
    ok = O.syndrome_ok(hard.ravel(), 10, deg, pos)
    assert list(ok) == [1, 0, 0]
