"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle must keep reproducing them (CPU), and the HIP decoder must
reproduce them bit for bit after every iteration (GPU)."""
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

FIXTURES = sorted((Path(__file__).parent / "golden").glob("*.npz"))
VARIANTS = [O.MS, O.OMS, O.OMS_V2, O.NMS, O.NMS2, O.LOMS, O.LNMS]


def table_for(f, v):
    return f["table_loms"] if v == O.LOMS else f["table_lnms"] if v == O.LNMS else None


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: p.stem)
@pytest.mark.parametrize("variant", VARIANTS)
def test_oracle_reproduces_golden(path, variant):
    f = np.load(path)
    n, W = int(f["n"]), int(f["W"])
    dec = O.OracleDecoder(f["deg"], f["pos"], n, W, variant, q1_split=int(f["q1_split"]), early_term=0,
                          table=table_for(f, variant))
    dec.load(f["llr"])
    for it in range(f[f"post_{variant}"].shape[0]):
        dec.iterate(1)
        np.testing.assert_array_equal(dec.post, f[f"post_{variant}"][it])
        np.testing.assert_array_equal(dec.msg, f[f"msg_{variant}"][it])
    np.testing.assert_array_equal(dec.hard(), f[f"hard_{variant}"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: p.stem)
@pytest.mark.parametrize("variant", VARIANTS)
def test_gpu_reproduces_golden(path, variant, gpu_available):
    import fast_ldpc_decoder_for_gpu_fixed_amd as L
    f = np.load(path)
    n, W = int(f["n"]), int(f["W"])
    code = L.Code.from_table(n, f["deg"], f["pos"], q1_split=int(f["q1_split"]))
    dec = L.make_decoder(code, W, variant, early_term=0)
    t = table_for(f, variant)
    if t is not None:
        dec.set_table(t)
    llr = f["llr"]
    dec.load(O.interleave(llr, n).view(np.int8) if variant == O.OMS_V2 else llr)
    for it in range(f[f"post_{variant}"].shape[0]):
        dec.iterate(1)
        post, msg = dec.get_state()
        np.testing.assert_array_equal(post, f[f"post_{variant}"][it])
        np.testing.assert_array_equal(msg, f[f"msg_{variant}"][it])
    np.testing.assert_array_equal(dec.get_hard(), f[f"hard_{variant}"])
