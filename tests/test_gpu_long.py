"""Long frames through the drop-in decoder: when one word block (4 frames,
N bytes each) does not fit in LDS, fixed-iteration decodes run on the
pair-mode frame queue with no stop test (LDPC_STOP_NONE).  Decisions equal
the oracle's fixed-iteration decode and the global-memory-posterior kernel's
for every variant that path takes; the per-word early stop keeps the global
kernel.  Also: the frame queue with LDPC_STOP_NONE runs every frame to the
limit and matches the fixed-iteration decoder."""
import numpy as np
import pytest

import fast_ldpc_decoder_for_gpu_fixed_amd as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N_LONG = 41000  # 4 * N bytes > the LDS budget of one word block


def llr_frames(n, frames, seed, ebn0=1.6):
    rng = np.random.default_rng(seed)
    sigma = np.sqrt(1.0 / (2 * 0.5 * 10 ** (ebn0 / 10)))
    y = -1.0 + sigma * rng.standard_normal((frames, n))
    return np.clip(np.round(16.0 * y / sigma ** 2), -127, 127).astype(np.int8)


_LONG = {}


def long_code():
    if "c" not in _LONG:
        _LONG["c"] = L.code_3_6(N_LONG, seed=5)
    return _LONG["c"]


def table():
    rng = np.random.default_rng(3)
    return rng.integers(0, 4, size=(12, 7)).astype(np.int8)


@pytest.mark.parametrize("variant", [L.MS, L.OMS, L.OMS_V2, L.NMS, L.NMS2, L.LOMS],
                         ids=lambda v: L.VARIANT_NAMES[v])
def test_long_frames_drop_in_matches_oracle(variant, gpu_available):
    code = long_code()
    deg, pos = code.export()
    W, iters = 16, 10
    llr = llr_frames(N_LONG, 4 * W, seed=variant)
    dec = L.make_decoder(code, W, variant, early_term=0)
    assert dec.config()["post_in_lds"] is False
    tab = table() if variant == L.LOMS else None
    if tab is not None:
        dec.set_table(tab)
    orc = O.OracleDecoder(deg, pos, N_LONG, W, variant, q1_split=code.q1_split, early_term=0, table=tab)
    # the xMS convention: OMS_v2 takes its input already in the device word layout
    x = O.interleave(llr, N_LONG).view(np.int8) if variant == L.OMS_V2 else llr
    out = dec.decode_bytes(x, iters)
    orc.load(llr)
    orc.iterate(iters, threads=8)
    np.testing.assert_array_equal(out, orc.hard())


def test_long_frames_equal_the_global_kernel(gpu_available):
    """The same frames through the forced global-memory-posterior kernel."""
    code = long_code()
    W = 16
    llr = llr_frames(N_LONG, 4 * W, seed=9)
    auto = L.make_decoder(code, W, L.NMS2, early_term=0).decode_bytes(llr, 12)
    forced = L.make_decoder(code, W, L.NMS2, early_term=0, post_in_lds=0).decode_bytes(llr, 12)
    np.testing.assert_array_equal(auto, forced)


@pytest.mark.parametrize("pairs", [0, 1], ids=["words", "pairs"])
def test_pool_stop_none_runs_to_the_limit(pairs, gpu_available):
    code = L.code_3_6(1000, seed=3)
    W = 64
    llr = llr_frames(1000, 4 * W, seed=2, ebn0=3.0)
    pool = L.PoolDecoder(code, L.OMS_V2, 9, stop_rule=L.STOP_NONE, pairs=pairs)
    hard, its = pool.decode(llr)
    assert (its == 9).all()
    ref = L.make_decoder(code, W, L.OMS_V2, early_term=0).decode_bytes(O.interleave(llr, 1000).view(np.int8), 9)
    np.testing.assert_array_equal(hard.reshape(-1), ref)
