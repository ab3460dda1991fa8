"""The mixed-code step on the quasi-cyclic workload it is benchmarked on.

Codes, channel and decoder exactly as `bench.py --mode mixed` builds them:
QC 12x24 Z=81 (seed 1, N=1944, rate 1/2, check degrees 3..9) and QC 4x24
Z=96 (seed 2, N=2304, rate 5/6, check degrees 10..16), frames interleaved
50/50 (frame f belongs to code f & 1, each code's frames contiguous in its own
region), per-frame Eb/N0 uniform in [1.5, 3.5] dB from the device channel
with the reference quantizer trunc(8y) clamped to +-31
(src/awgn_channel/CChanel_AWGN_SIMD.cu:17-28), OMS with the clamp on every
check (OMS_v2), at most 30 iterations, per-frame syndrome stop.  The
reference decodes codes with two check degrees with its multi-degree kernel
(src/decoder_oms_v2/cuda/CUDA_OMS_SIMD_v2.cu:343-575, selected at
src/decoder_oms_v2/CGPU_Decoder_MS_SIMD_v2.cu:222-238).

Every frame's decisions and iteration count must equal the oracle's per-frame
decode of that frame with its own code: with the autotuned tuning of each code
(eager, and hipGraph replays with eager work between them), and with each
tuning the autotuner chooses among (4-frame words at P = 1, 2, 4; 2-frame
pairs at P = 2, 4, 8), since its pick varies from run to run with timing."""
import os

import numpy as np
import pytest

import fast_ldpc_decoder_for_gpu_fixed_amd as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
FRAMES = 4096  # the first 2048 frames of each code's channel stream in the benchmark
MAX_IT = 30
SEED = 2024

_W = {}


def workload():
    if not _W:
        codes = [L.Code.qc(12, 24, 81, 0.35, seed=1), L.Code.qc(4, 24, 96, 0.35, seed=2)]
        F = FRAMES
        cid = (np.arange(F) & 1).astype(np.uint8)
        n_c = [(F + 1 - k) // 2 for k in range(2)]
        base = [0, n_c[0] * codes[0].n_vars]
        off = np.empty(F, np.uint64)
        for k in range(2):
            off[k::2] = base[k] + np.arange(n_c[k], dtype=np.uint64) * codes[k].n_vars
        total = base[1] + n_c[1] * codes[1].n_vars
        d_in = L.DeviceBuffer(total)
        for k, c in enumerate(codes):
            L.channel_generate_range(d_in.ptr + base[k], c.n_vars, n_c[k], SEED + k, 1.5, 3.5, c.rate, quant=1)
        L.synchronize()
        buf = d_in.download(np.empty(total, np.int8))
        # the oracle: each code's frames with their own code
        ref_h = np.zeros(total, np.uint8)
        ref_i = np.zeros(F, np.int32)
        for k, c in enumerate(codes):
            deg, pos = c.export()
            llr = buf[base[k]:base[k] + n_c[k] * c.n_vars].reshape(n_c[k], c.n_vars)
            h, i = O.decode_frames(deg, pos, c.n_vars, O.OMS_V2, llr, MAX_IT, threads=THREADS)
            ref_h[base[k]:base[k] + n_c[k] * c.n_vars] = h.reshape(-1)
            ref_i[k::2] = i
        d_off, d_cid = L.DeviceBuffer(off.nbytes), L.DeviceBuffer(F)
        d_off.upload(off)
        d_cid.upload(cid)
        _W.update(codes=codes, cid=cid, off=off, total=total, base=base, n_c=n_c, d_in=d_in, d_off=d_off,
                  d_cid=d_cid, ref_h=ref_h, ref_i=ref_i, d_out=L.DeviceBuffer(total), d_it=L.DeviceBuffer(4 * F))
    return _W


def run_step(w, mixed, graph):
    w["d_out"].upload(np.full(w["total"], 0xAA, np.uint8))
    w["d_it"].upload(np.full(FRAMES, -1, np.int32))
    mixed.decode_device(w["d_in"].ptr, w["d_off"].ptr, w["d_cid"].ptr, FRAMES, w["d_out"].ptr, w["d_it"].ptr,
                        graph=graph)
    L.synchronize()
    return (w["d_out"].download(np.empty(w["total"], np.uint8)), w["d_it"].download(np.empty(FRAMES, np.int32)))


def check_equal(w, hard, its, what):
    for k in range(2):
        bad = np.nonzero(its[k::2] != w["ref_i"][k::2])[0]
        assert bad.size == 0, f"{what}: code {k}: {bad.size} iteration counts differ, frames {(2 * bad + k)[:8]}"
    bad = np.nonzero(hard != w["ref_h"])[0]
    assert bad.size == 0, f"{what}: {bad.size} decision bytes differ, first at byte {bad[:4].tolist()}"


def test_workload_shape(gpu_available):
    """The two codes' structure and iteration profile as the benchmark reports them."""
    w = workload()
    c0, c1 = w["codes"]
    assert (c0.n_vars, c0.n_checks, c1.n_vars, c1.n_checks) == (1944, 972, 2304, 384)
    deg0, _ = c0.export()
    deg1, _ = c1.export()
    assert len(set(deg0.tolist())) > 2 and len(set(deg1.tolist())) > 2  # several check degrees each
    assert 8 < c1.max_deg <= 16 and c0.max_deg <= 10
    i0, i1 = w["ref_i"][0::2], w["ref_i"][1::2]
    assert len(set(i0.tolist())) > 5 and i0.mean() < 15  # uneven, early stop on the rate-1/2 code
    assert i1.mean() > 25  # the rate-5/6 code mostly runs to the limit in [1.5, 3.5] dB


def test_c3_autotuned_eager_and_graph(gpu_available):
    w = workload()
    tuned = [L.pool_autotune(c, L.OMS_V2, MAX_IT, ebn0_db=2.5, sample_frames=32768, quant=1) for c in w["codes"]]
    print("autotuned:", tuned, flush=True)
    mixed = L.MixedDecoder(w["codes"], L.OMS_V2, MAX_IT, tunings=tuned, max_frames=FRAMES)
    for k, t in enumerate(tuned):
        cfg = mixed.config(k)
        assert (cfg["words_per_block"], cfg["threads"], cfg["pairs"]) == (t["words_per_block"], t["threads"],
                                                                          t["pairs"])
    hard, its = run_step(w, mixed, graph=False)
    check_equal(w, hard, its, "eager")
    for rep in range(3):
        hard, its = run_step(w, mixed, graph=True)
        check_equal(w, hard, its, f"graph replay {rep}")
        h2, i2 = run_step(w, mixed, graph=False)  # eager work between replays
        check_equal(w, h2, i2, f"eager after replay {rep}")


TUNINGS = [
    ("words_P1_T64", {"pairs": 0, "words_per_block": 1, "threads": 64}),
    ("words_P1_T128", {"pairs": 0, "words_per_block": 1, "threads": 128}),
    ("words_P2_T128", {"pairs": 0, "words_per_block": 2, "threads": 128}),
    ("words_P2_T256", {"pairs": 0, "words_per_block": 2, "threads": 256}),
    ("words_P4_T256", {"pairs": 0, "words_per_block": 4, "threads": 256}),
    ("pairs_P2_T256", {"pairs": 1, "words_per_block": 2, "threads": 256}),
    ("pairs_P4_T256", {"pairs": 1, "words_per_block": 4, "threads": 256}),
    ("pairs_P8_T256", {"pairs": 1, "words_per_block": 8, "threads": 256}),
]


@pytest.mark.parametrize("name,tuning", TUNINGS, ids=[t[0] for t in TUNINGS])
def test_c3_each_candidate_tuning(name, tuning, gpu_available):
    w = workload()
    mixed = L.MixedDecoder(w["codes"], L.OMS_V2, MAX_IT, tunings=[tuning, tuning], max_frames=FRAMES)
    for graph in (False, True):
        hard, its = run_step(w, mixed, graph=graph)
        check_equal(w, hard, its, f"{name} graph={graph}")
