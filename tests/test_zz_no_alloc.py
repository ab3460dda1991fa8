"""ldpc_mixed_decode_noalloc: the mixed-code step with its scratch (queue
counters and per-code frame lists) in the library's static device array.

Called through the C ABI, its decisions and iteration counts must equal
ldpc_mixed_decode's (eager) on the same inputs, bit for bit, on calls in a
row whose inputs shrink and grow again (the array keeps the last call's
contents; the step clears what must start at zero), and sizes beyond the
array are rejected before anything is enqueued."""
import ctypes as C

import numpy as np
import pytest

import fast_ldpc_decoder_for_gpu_fixed_amd as L

pytestmark = pytest.mark.gpu


def workload(cs, F, seed):
    rng = np.random.default_rng(seed)
    cid = rng.integers(0, len(cs), size=F).astype(np.uint8)
    off, total = L.mixed_layout(cs, cid)
    buf = np.empty(total, np.int8)
    for f in range(F):
        c = cs[cid[f]]
        sigma = np.sqrt(1.0 / (2 * c.rate * 10 ** (rng.uniform(1.5, 4.0) / 10)))
        y = -1.0 + sigma * rng.standard_normal(c.n_vars)
        buf[off[f]:off[f] + c.n_vars] = np.trunc(np.clip(8.0 * y, -31, 31))
    return cid, off, buf


class Step:
    def __init__(self, cid, off, buf):
        self.F, self.nbytes = len(cid), buf.nbytes
        self.d_in, self.d_off, self.d_cid = L.DeviceBuffer(buf.nbytes), L.DeviceBuffer(off.nbytes), L.DeviceBuffer(self.F)
        self.d_in.upload(buf)
        self.d_off.upload(off)
        self.d_cid.upload(cid)
        self.d_out, self.d_it = L.DeviceBuffer(buf.nbytes), L.DeviceBuffer(4 * self.F)

    def result(self):
        L.synchronize()
        return (self.d_out.download(np.empty(self.nbytes, np.uint8)), self.d_it.download(np.empty(self.F, np.int32)))

    def clobber(self):
        self.d_out.upload(np.full(self.nbytes, 0xAA, np.uint8))
        self.d_it.upload(np.full(self.F, -1, np.int32))

    def reference(self, mixed):
        self.clobber()
        mixed.decode_device(self.d_in.ptr, self.d_off.ptr, self.d_cid.ptr, self.F, self.d_out.ptr, self.d_it.ptr)
        return self.result()

    def noalloc(self, mixed):
        self.clobber()
        rc = L.load().ldpc_mixed_decode_noalloc(mixed._h, C.c_void_p(self.d_in.ptr), C.c_void_p(self.d_off.ptr),
                                                C.c_void_p(self.d_cid.ptr), self.F, C.c_void_p(self.d_out.ptr),
                                                C.c_void_p(self.d_it.ptr), None)
        L.check(rc, "ldpc_mixed_decode_noalloc")
        return self.result()


@pytest.mark.parametrize("variant", [L.OMS_V2, L.NMS2], ids=lambda v: L.VARIANT_NAMES[v])
def test_noalloc_step_equals_the_allocating_step(variant, gpu_available):
    cs = [L.Code.qc(12, 24, 27, 0.35, seed=3), L.Code.qc(4, 24, 24, 0.35, seed=4), L.code_3_6(1000, seed=3)]
    mixed = L.MixedDecoder(cs, variant, 20)
    big, small = Step(*workload(cs, 3000, seed=1)), Step(*workload(cs, 1100, seed=2))
    for name, step in (("first call, 3000 frames", big), ("second call, 1100 frames", small),
                       ("third call, 3000 frames again", big)):
        print(name, L.VARIANT_NAMES[variant], flush=True)
        got_h, got_i = step.noalloc(mixed)
        ref_h, ref_i = step.reference(mixed)
        assert (ref_i > 0).all()
        np.testing.assert_array_equal(got_i, ref_i, err_msg=f"{name}: iterations")
        np.testing.assert_array_equal(got_h, ref_h, err_msg=f"{name}: decisions")


def test_noalloc_rejects_sizes_beyond_the_scratch(gpu_available):
    cs = [L.code_3_6(1000, seed=3), L.Code.qc(4, 24, 24, 0.35, seed=4)]
    mixed = L.MixedDecoder(cs, L.OMS_V2, 10)
    step = Step(*workload(cs, 64, seed=3))
    too_many = L.MixedDecoder.NOALLOC_ENTRIES // len(cs) + 1
    # rejected on the host before anything is enqueued: no buffer is read
    rc = L.load().ldpc_mixed_decode_noalloc(mixed._h, C.c_void_p(step.d_in.ptr), C.c_void_p(step.d_off.ptr),
                                            C.c_void_p(step.d_cid.ptr), too_many, C.c_void_p(step.d_out.ptr),
                                            C.c_void_p(step.d_it.ptr), None)
    assert rc != 0 and "static scratch" in L._lib.last_error()
    # the largest accepted size is frames * codes == the entry count
    assert (L.MixedDecoder.NOALLOC_ENTRIES // len(cs)) * len(cs) <= L.MixedDecoder.NOALLOC_ENTRIES
    got_h, got_i = step.noalloc(mixed)
    ref_h, ref_i = step.reference(mixed)
    np.testing.assert_array_equal(got_i, ref_i)
    np.testing.assert_array_equal(got_h, ref_h)
