"""The CPU reference configuration through the drop-in GPU decoder: the
(3,6) N=4000 code, plain OMS (offset 1), 20 fixed iterations, 4096 frames
(1024 words) at Eb/N0 = 1.0 .. 3.0 dB in 0.5 dB steps.  At every point the
GPU's hard decisions equal the oracle's on the same quantised LLRs, and the
frame error count falls along the sweep (the waterfall)."""
import numpy as np
import pytest

import fast_ldpc_decoder_for_gpu_fixed_amd as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu

EBN0 = (1.0, 1.5, 2.0, 2.5, 3.0)


def test_reference_sweep_bit_exact_and_waterfall(gpu_available):
    code = L.code_3_6(4000, seed=1)
    deg, pos = code.export()
    N, W = 4000, 1024
    F = 4 * W
    dec = L.make_decoder(code, W, L.OMS, early_term=0)
    frame_errors = []
    for i, ebn0 in enumerate(EBN0):
        llr = O.channel_frames(N, 0, F, 42, i, O.sigma(ebn0, code.rate), 0)
        print(f"sweep point {ebn0} dB", flush=True)
        out = dec.decode_bytes(llr, 20)
        orc = O.OracleDecoder(deg, pos, N, W, O.OMS, q1_split=code.q1_split, early_term=0)
        orc.load(llr)
        orc.iterate(20, threads=8)
        np.testing.assert_array_equal(out, orc.hard())
        fe, _, _ = O.count_errors(out, N, code.n_info)
        frame_errors.append(fe)
    # 1.0 dB is below the code's threshold, 3.0 dB well above it
    assert frame_errors[0] > 0.5 * F, frame_errors
    assert frame_errors[-1] <= frame_errors[2] < frame_errors[0], frame_errors
    assert all(a >= b for a, b in zip(frame_errors, frame_errors[1:])), frame_errors
