"""Pins of the oracle's per-frame stop decode (the model of the persistent
frame-queue decoder).

v2c-parity rule: checked against the word decoder with the reference's
early-termination rule (src/decoder_2nms/cuda/CUDA_2NMS_SIMD.cu:237-244): a
frame copied into all four byte lanes of a word makes the word's parity the
frame's own, and starting that word with zero messages but `first` cleared
applies the stop test after the first iteration too (t = post - 0 = post is
the first iteration's value).  Both paths must give the same iteration count
and decisions for every frame.

Syndrome rule: a frame stopped after k iterations has decisions equal to a
k-iteration fixed decode of that frame, those decisions satisfy every check
(numpy parity of H times the decisions), and after k - 1 iterations they did
not (unless k hit the limit).  A noiseless frame stops after one iteration.
"""
import numpy as np
import pytest

from oracle import oracle as O

VARIANTS = [O.MS, O.OMS_V2, O.NMS, O.NMS2, O.LOMS, O.LNMS]


def _code(n=480, seed=2):
    import fast_ldpc_decoder_for_gpu_fixed_amd as L
    c = L.code_3_6(n, seed=seed)
    deg, pos = c.export()
    return c, deg, pos


def _llr(n, frames, seed, lo=1.0, hi=3.5):
    rng = np.random.default_rng(seed)
    ebn0 = rng.uniform(lo, hi, size=(frames, 1))
    sigma = np.sqrt(1.0 / (2 * 0.5 * 10 ** (ebn0 / 10)))
    y = -1.0 + sigma * rng.standard_normal((frames, n))
    return np.clip(np.round(16.0 * y / sigma ** 2), -127, 127).astype(np.int8)


def _table(variant, iters=12, degs=7, seed=3):
    if variant not in (O.LOMS, O.LNMS):
        return None
    rng = np.random.default_rng(seed)
    lo, hi = (0, 3) if variant == O.LOMS else (5, 8)
    return rng.integers(lo, hi + 1, size=(iters, degs)).astype(np.int8)


def np_syndrome_ok(hard, deg, pos):
    """Every check's XOR of decisions is 0 (plain numpy, per frame)."""
    hard = np.asarray(hard).reshape(len(hard), -1)
    x = hard[:, pos].astype(np.int64)
    bounds = np.concatenate([[0], np.cumsum(deg)])[:-1]
    return (np.add.reduceat(x, bounds, axis=1) % 2 == 0).all(axis=1)


@pytest.mark.parametrize("variant", VARIANTS)
def test_frames_syndrome_rule(variant):
    code, deg, pos = _code()
    n, F, max_it = code.n_vars, 48, 12
    llr = _llr(n, F, seed=30 + variant)
    table = _table(variant)
    hard, its = O.decode_frames(deg, pos, n, variant, llr, max_it, table=table, threads=4,
                                stop_rule=O.STOP_SYNDROME)
    ok = np_syndrome_ok(hard, deg, pos)
    assert ok[its < max_it].all()
    rep = np.repeat(llr, 4, axis=0).reshape(-1)
    for k in sorted(set(its.tolist())):
        sel = np.nonzero(its == k)[0]
        for kk, want in ((k, None), (k - 1, False)):
            if kk < 1:
                continue
            d = O.OracleDecoder(deg, pos, n, F, variant, early_term=0, table=table)
            d.load(rep)
            d.iterate(kk)
            h = d.hard().reshape(F, 4, n)[:, 0]
            if want is None:
                np.testing.assert_array_equal(h[sel], hard[sel])
            else:
                assert not np_syndrome_ok(h[sel], deg, pos).any(), f"stopped late at {k}"
    assert (its < max_it).any() and len(set(its.tolist())) > 2


@pytest.mark.parametrize("variant", VARIANTS)
def test_frames_match_replicated_word_early_term(variant):
    code, deg, pos = _code()
    n, F, max_it = code.n_vars, 96, 12
    llr = _llr(n, F, seed=10 + variant)
    table = _table(variant)
    hard, its = O.decode_frames(deg, pos, n, variant, llr, max_it, table=table, threads=4,
                                stop_rule=O.STOP_V2C_PARITY)

    rep = np.repeat(llr, 4, axis=0)  # word w = frame w in all four lanes
    ref = O.OracleDecoder(deg, pos, n, F, variant, early_term=1, table=table)
    ref.load(rep.reshape(-1))
    ref.first = False  # zero messages: the first iteration is tested too
    ref.iterate(max_it)
    np.testing.assert_array_equal(its, ref.iters_run)
    np.testing.assert_array_equal(hard, ref.hard().reshape(F, 4, n)[:, 0])
    # the mix has frames stopping early, late and not at all
    assert (its < max_it).any() and (its == max_it).any()


def test_noiseless_frame_stops_after_one_iteration():
    code, deg, pos = _code()
    llr = np.full((5, code.n_vars), -20, np.int8)
    hard, its = O.decode_frames(deg, pos, code.n_vars, O.OMS_V2, llr, 10)
    assert (its == 1).all() and not hard.any()


def test_frames_are_independent_of_their_neighbours():
    code, deg, pos = _code()
    llr = _llr(code.n_vars, 7, seed=4)
    hard, its = O.decode_frames(deg, pos, code.n_vars, O.MS, llr, 15)
    for f in range(7):
        h1, i1 = O.decode_frames(deg, pos, code.n_vars, O.MS, llr[f:f + 1], 15)
        assert i1[0] == its[f]
        np.testing.assert_array_equal(h1[0], hard[f])


def test_frames_refuse_oms_group_rule():
    code, deg, pos = _code()
    with pytest.raises(ValueError):
        O.decode_frames(deg, pos, code.n_vars, O.OMS, np.zeros((1, code.n_vars), np.int8), 3)
