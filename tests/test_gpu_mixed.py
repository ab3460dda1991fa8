"""GPU parity of the mixed-code decode step (ldpc_mixed_*) and the autotuner.

Frames of several codes sit back to back in one buffer with a code id and a
byte offset each; every frame's decisions and iteration count must equal the
oracle's per-frame stop decode of that frame with its own code.  The hipGraph
path is replayed several times with eager work and synchronisations between
the replays, and every replay is checked against the eager result."""
import os

import numpy as np
import pytest

import fast_ldpc_decoder_for_gpu_fixed_amd as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)

_C = {}


def codes():
    if not _C:
        _C["list"] = [L.Code.qc(12, 24, 27, 0.35, seed=3), L.Code.qc(4, 24, 24, 0.35, seed=4),
                      L.code_3_6(1000, seed=3)]
    return _C["list"]


def workload(cs, F, seed):
    rng = np.random.default_rng(seed)
    cid = rng.integers(0, len(cs), size=F).astype(np.uint8)
    off, total = L.mixed_layout(cs, cid)
    buf = np.empty(total, np.int8)
    for f in range(F):
        c = cs[cid[f]]
        rate = c.n_info / c.n_vars
        eb = rng.uniform(1.5, 4.5)
        sigma = np.sqrt(1.0 / (2 * rate * 10 ** (eb / 10)))
        y = -1.0 + sigma * rng.standard_normal(c.n_vars)
        buf[off[f]:off[f] + c.n_vars] = np.clip(np.round(16.0 * y / sigma ** 2), -127, 127)
    return cid, off, buf


def oracle_mixed(cs, cid, off, buf, variant, max_it, tables=None):
    F = len(cid)
    hard = np.zeros(buf.size, np.uint8)
    its = np.zeros(F, np.int32)
    for k, c in enumerate(cs):
        idx = np.nonzero(cid == k)[0]
        if idx.size == 0:
            continue
        llr = np.stack([buf[off[f]:off[f] + c.n_vars] for f in idx])
        deg, pos = c.export()
        h, i = O.decode_frames(deg, pos, c.n_vars, variant, llr, max_it,
                               table=None if tables is None else tables[k], threads=THREADS)
        for j, f in enumerate(idx):
            hard[off[f]:off[f] + c.n_vars] = h[j]
        its[idx] = i
    return hard, its


class Bufs:
    def __init__(self, cid, off, buf):
        self.F = len(cid)
        self.d_in = L.DeviceBuffer(buf.nbytes)
        self.d_in.upload(buf)
        self.d_off = L.DeviceBuffer(off.nbytes)
        self.d_off.upload(off)
        self.d_cid = L.DeviceBuffer(cid.nbytes)
        self.d_cid.upload(cid)
        self.d_out = L.DeviceBuffer(buf.nbytes)
        self.d_it = L.DeviceBuffer(4 * self.F)
        self.nbytes = buf.nbytes

    def run(self, mixed, graph=False, out=None, it=None):
        out = out or self.d_out
        it = it or self.d_it
        mixed.decode_device(self.d_in.ptr, self.d_off.ptr, self.d_cid.ptr, self.F, out.ptr, it.ptr, graph=graph)
        L.synchronize()
        return out.download(np.empty(self.nbytes, np.uint8)), it.download(np.empty(self.F, np.int32))


@pytest.mark.parametrize("variant", [L.OMS_V2, L.NMS2, L.LOMS], ids=lambda v: L.VARIANT_NAMES[v])
def test_mixed_parity(variant, gpu_available):
    cs = codes()
    cid, off, buf = workload(cs, 1500, seed=variant)
    tables = None
    mixed = L.MixedDecoder(cs, variant, 25)
    if variant == L.LOMS:
        rng = np.random.default_rng(9)
        tables = [rng.integers(0, 3, size=(25, 33)).astype(np.int8) for _ in cs]
        for k, t in enumerate(tables):
            mixed.set_table(k, t)
    b = Bufs(cid, off, buf)
    hard, its = b.run(mixed)
    ref_h, ref_i = oracle_mixed(cs, cid, off, buf, variant, 25, tables)
    np.testing.assert_array_equal(its, ref_i, err_msg="iterations per frame")
    np.testing.assert_array_equal(hard, ref_h, err_msg="decisions")
    assert len(set(its.tolist())) > 3  # uneven iteration counts


def test_mixed_graph_replays_match_eager(gpu_available):
    cs = codes()
    cid, off, buf = workload(cs, 1200, seed=21)
    mixed = L.MixedDecoder(cs, L.OMS_V2, 20, max_frames=4096)
    b = Bufs(cid, off, buf)
    eager_h, eager_i = b.run(mixed)
    ref_h, ref_i = oracle_mixed(cs, cid, off, buf, L.OMS_V2, 20)
    np.testing.assert_array_equal(eager_i, ref_i)
    np.testing.assert_array_equal(eager_h, ref_h)
    other_out, other_it = L.DeviceBuffer(b.nbytes), L.DeviceBuffer(4 * b.F)
    for rep in range(4):
        # clobber the outputs, replay, and do eager work between replays
        b.d_out.upload(np.full(b.nbytes, 0xAA, np.uint8))
        b.d_it.upload(np.full(b.F, -1, np.int32))
        h, i = b.run(mixed, graph=True)
        np.testing.assert_array_equal(i, eager_i, err_msg=f"graph replay {rep}")
        np.testing.assert_array_equal(h, eager_h, err_msg=f"graph replay {rep}")
        h2, i2 = b.run(mixed, graph=False, out=other_out, it=other_it)
        np.testing.assert_array_equal(i2, eager_i)
    # other buffers: the step is captured again
    h, i = b.run(mixed, graph=True, out=other_out, it=other_it)
    np.testing.assert_array_equal(i, eager_i)
    np.testing.assert_array_equal(h, eager_h)


def test_mixed_unknown_code_id_is_skipped(gpu_available):
    cs = codes()[:2]
    cid, off, buf = workload(cs, 300, seed=5)
    cid = cid.copy()
    cid[7] = 9
    mixed = L.MixedDecoder(cs, L.MS, 15)
    b = Bufs(cid, off, buf)
    b.d_out.upload(np.full(b.nbytes, 0xAA, np.uint8))
    hard, its = b.run(mixed)
    assert its[7] == 0
    n7 = int(off[8] - off[7])  # the frame's length in the layout
    assert (hard[off[7]:off[7] + n7] == 0xAA).all()
    keep = cid < len(cs)
    ref_h, ref_i = oracle_mixed(cs, np.where(keep, cid, 255).astype(np.uint8), off, buf, L.MS, 15)
    np.testing.assert_array_equal(its[keep], ref_i[keep])
    with pytest.raises(L.LdpcError):  # an empty step
        mixed.decode_device(b.d_in.ptr, b.d_off.ptr, b.d_cid.ptr, 0, b.d_out.ptr)


def test_autotune_picks_a_working_configuration(gpu_available):
    c = codes()[2]
    best = L.pool_autotune(c, L.OMS_V2, 20, ebn0_db=2.5, sample_frames=4096)
    assert best["tried"] > 4 and best["ms"] > 0
    assert best["uniform"] in (0, 1) and best["words_per_block"] >= 1 and best["threads"] >= 64
    pool = L.PoolDecoder(c, L.OMS_V2, 20, words_per_block=best["words_per_block"], threads=best["threads"],
                         uniform=best["uniform"], batch=best["batch"])
    rng = np.random.default_rng(2)
    sigma = np.sqrt(1.0 / (2 * 0.5 * 10 ** 0.2))
    llr = np.clip(np.round(16.0 * (-1.0 + sigma * rng.standard_normal((800, c.n_vars))) / sigma ** 2),
                  -127, 127).astype(np.int8)
    hard, its = pool.decode(llr)
    deg, pos = c.export()
    rh, ri = O.decode_frames(deg, pos, c.n_vars, L.OMS_V2, llr, 20, threads=THREADS)
    np.testing.assert_array_equal(its, ri)
    np.testing.assert_array_equal(hard, rh)
    # and the mixed decoder accepts tuned configurations per code
    tuned = [L.pool_autotune(cc, L.OMS_V2, 20, sample_frames=2048) for cc in codes()]
    mixed = L.MixedDecoder(codes(), L.OMS_V2, 20, tunings=tuned)
    for k, t in enumerate(tuned):
        cfg = mixed.config(k)
        assert cfg["words_per_block"] == t["words_per_block"] and cfg["threads"] == t["threads"]


def test_mixed_with_pair_mode_code(gpu_available):
    """One code in pair mode (2 frames per 16-bit entry), the others in words,
    in one step; eager and graph both equal the oracle."""
    cs = codes()
    cid, off, buf = workload(cs, 1000, seed=77)
    tunings = [{"pairs": 1, "words_per_block": 2, "threads": 256}, None, {"pairs": 0}]
    mixed = L.MixedDecoder(cs, L.OMS_V2, 20, tunings=tunings)
    assert mixed.config(0)["pairs"] == 1 and mixed.config(2)["pairs"] == 0
    b = Bufs(cid, off, buf)
    ref_h, ref_i = oracle_mixed(cs, cid, off, buf, L.OMS_V2, 20)
    for graph in (False, True, True):
        hard, its = b.run(mixed, graph=graph)
        np.testing.assert_array_equal(its, ref_i)
        np.testing.assert_array_equal(hard, ref_h)
