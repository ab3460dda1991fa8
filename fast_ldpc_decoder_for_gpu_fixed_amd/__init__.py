"""MI355X-native 8-bit fixed-point Min-Sum LDPC decoder (gfx950).

Python face of ``libldpc_mi355x.so`` (C-ABI: ``include/ldpc_mi355x.h``).  It
mirrors the reference's decoder objects (``src/decoder_template/CGPUDecoder.h``
and the five subclasses) so that tests and drivers read like the reference:

    dec = CGPU_Decoder_OMS_SIMD(nb_frames=W, n=N, k=checks, m=edges, code=code)
    dec.initialize()
    dec.decode(noise, decoded, 20)

All decoding runs in the HIP library; nothing here computes a decode on the
CPU, and every entry point raises if the library is missing.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from ._lib import LdpcError, LearnConfig, SimStats, Tuning, check, load, ptr

MS, OMS, OMS_V2, NMS, NMS2, LOMS, LNMS = range(7)
STOP_SYNDROME, STOP_V2C_PARITY, STOP_NONE = 0, 1, 2  # frame-queue decoder stop tests
ORDER_INPUT, ORDER_LEAST_RELIABLE_FIRST = 0, 1  # frame-queue decoder queue orders
VARIANT_NAMES = {MS: "MS", OMS: "OMS", OMS_V2: "OMS_v2", NMS: "NMS", NMS2: "2NMS",
                 LOMS: "learned-OMS", LNMS: "learned-NMS"}


# ---------------------------------------------------------------------------
# codes (replaces the compile-time tables of src/matrix)
# ---------------------------------------------------------------------------
class Code:
    """A parity-check table: checks in processing order, each with its variables."""

    def __init__(self, handle: int):
        if not handle:
            raise LdpcError(f"code creation failed: {_lib.last_error()}")
        self._h = C.c_void_p(handle)
        n, c, e, d, lv, q = (C.c_int() for _ in range(6))
        check(load().ldpc_code_info(self._h, C.byref(n), C.byref(c), C.byref(e), C.byref(d),
                                    C.byref(lv), C.byref(q)), "ldpc_code_info")
        self.n_vars, self.n_checks, self.n_edges = n.value, c.value, e.value
        self.max_deg, self.n_levels, self.q1_split = d.value, lv.value, q.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.ldpc_code_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @property
    def n_info(self) -> int:
        """N - #checks: the positions the reference's error counter scans
        (src/ber_analyzer/CErrorAnalyzer.cpp:147, _vars = N - K)."""
        return self.n_vars - self.n_checks

    @property
    def rate(self) -> float:
        return self.n_info / self.n_vars

    @classmethod
    def from_table(cls, n_vars: int, deg, pos, q1_split: int = -1) -> "Code":
        deg = np.ascontiguousarray(deg, dtype=np.int32)
        pos = np.ascontiguousarray(pos, dtype=np.uint32)
        if int(deg.sum()) != pos.size:
            raise LdpcError("sum(deg) != len(pos)")
        return cls(load().ldpc_code_create(n_vars, deg.size, ptr(deg, C.c_int), ptr(pos, C.c_uint32), q1_split))

    @classmethod
    def grouped(cls, n_vars, deg1, count1, deg2, count2, pos) -> "Code":
        pos = np.ascontiguousarray(pos, dtype=np.uint32)
        return cls(load().ldpc_code_create_grouped(n_vars, deg1, count1, deg2, count2, ptr(pos, C.c_uint32)))

    @classmethod
    def load_table(cls, path: str, n_vars: int = 0) -> "Code":
        return cls(load().ldpc_code_load_table(str(path).encode(), n_vars))

    @classmethod
    def regular(cls, n_vars: int, dv: int, dc: int, seed: int = 1) -> "Code":
        return cls(load().ldpc_code_regular(n_vars, dv, dc, seed))

    @classmethod
    def qc(cls, mb: int, nb: int, z: int, p: float, seed: int = 1) -> "Code":
        return cls(load().ldpc_code_qc(mb, nb, z, p, seed))

    @classmethod
    def ira(cls, k_info: int, n_hi: int, d_hi: int, d_lo: int, seed: int = 1) -> "Code":
        return cls(load().ldpc_code_ira(k_info, n_hi, d_hi, d_lo, seed))

    def reorder_by_colour(self) -> "Code":
        return Code(load().ldpc_code_reorder_by_colour(self._h))

    def export(self):
        deg = np.empty(self.n_checks, np.int32)
        pos = np.empty(self.n_edges, np.uint32)
        check(load().ldpc_code_export(self._h, ptr(deg, C.c_int), ptr(pos, C.c_uint32)), "ldpc_code_export")
        return deg, pos

    def levels(self, granule: int = 1):
        """Level of every check in the decoders' parallel schedule (granule > 1:
        level sizes in multiples of granule where the dependencies allow)."""
        out = np.empty(self.n_checks, np.int32)
        n = load().ldpc_code_levels(self._h, granule, ptr(out, C.c_int))
        if n < 0:
            check(n, "ldpc_code_levels")
        return out


def code_3_6(n: int = 4000, seed: int = 1, colour: bool = True) -> Code:
    """(dv=3, dc=6) code from a seeded socket permutation with 4-cycles rejected;
    by default its table lists checks colour class by colour class."""
    c = Code.regular(n, 3, 6, seed)
    return c.reorder_by_colour() if colour else c


# ---------------------------------------------------------------------------
# decoders (replaces class CGPUDecoder and subclasses)
# ---------------------------------------------------------------------------
class CGPUDecoder:
    """src/decoder_template/CGPUDecoder.h:19-35 on the HIP library.

    nb_frames is a count of 32-bit words (4 frames each), as in the reference
    (src/main.cpp:210); n, k, m are variables, checks, edges and must match
    the code."""

    VARIANT: Optional[int] = None

    def __init__(self, nb_frames: int, n: int, k: int, m: int, code: Code,
                 words_per_block: int = 0, threads: int = 0, early_term: int = -1,
                 post_in_lds: int = -1, uniform: int = -1, batch: int = 0,
                 variant: Optional[int] = None):
        if (n, k, m) != (code.n_vars, code.n_checks, code.n_edges):
            raise LdpcError(f"(n, k, m) = {(n, k, m)} does not match the code "
                            f"{(code.n_vars, code.n_checks, code.n_edges)}")
        self.variant = self.VARIANT if variant is None else variant
        if self.variant is None:
            raise LdpcError("CGPUDecoder is abstract; use a variant subclass")
        self.code = code
        self.nb_frames = nb_frames
        self.n, self.k, self.m = n, k, m
        tu = Tuning(words_per_block, threads, early_term, post_in_lds, uniform, batch)
        h = load().ldpc_decoder_create(code.handle, nb_frames, self.variant, C.byref(tu))
        if not h:
            raise LdpcError(f"ldpc_decoder_create failed: {_lib.last_error()}")
        self._h = C.c_void_p(h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.ldpc_decoder_destroy(h)
            self._h = None

    @property
    def frames(self) -> int:
        return 4 * self.nb_frames

    @property
    def io_bytes(self) -> int:
        return self.frames * self.n

    def initialize(self) -> None:
        check(load().ldpc_decoder_initialize(self._h), "initialize")

    def _in(self, var_nodes) -> np.ndarray:
        a = np.ascontiguousarray(var_nodes).view(np.int8).reshape(-1)
        if a.size != self.io_bytes:
            raise LdpcError(f"input holds {a.size} bytes, expected {self.io_bytes}")
        return a

    def decode(self, var_nodes, Rprime_fix, nombre_iterations: int) -> None:
        """decode(float var_nodes[], int Rprime_fix[], int iterations): the
        buffers are reinterpreted as bytes, as in the reference."""
        a = self._in(var_nodes)
        out = Rprime_fix.view(np.uint8).reshape(-1)
        if out.size != self.io_bytes or not Rprime_fix.flags.c_contiguous:
            raise LdpcError("output buffer must be contiguous with 4W*N bytes")
        check(load().ldpc_decoder_decode(self._h, ptr(a, C.c_int8), ptr(out, C.c_uint8), nombre_iterations), "decode")

    def decode_bytes(self, llr, iterations: int) -> np.ndarray:
        out = np.empty(self.io_bytes, np.uint8)
        self.decode(llr, out, iterations)
        return out

    def decode_device(self, d_in: int, d_out: int, iterations: int, stream: int = 0) -> None:
        check(load().ldpc_decoder_decode_device(self._h, C.c_void_p(d_in), C.c_void_p(d_out), iterations,
                                                C.c_void_p(stream) if stream else None), "decode_device")

    @property
    def stream(self) -> int:
        return int(load().ldpc_decoder_stream(self._h) or 0)

    def config(self) -> dict:
        p, t, l, lv, u, b = (C.c_int() for _ in range(6))
        check(load().ldpc_decoder_config(self._h, C.byref(p), C.byref(t), C.byref(l), C.byref(lv), C.byref(u),
                                         C.byref(b)), "config")
        return {"words_per_block": p.value, "threads": t.value, "post_in_lds": bool(l.value),
                "n_levels": lv.value, "uniform_deg": u.value, "batch": b.value}

    def set_table(self, table) -> None:
        t = np.ascontiguousarray(table, dtype=np.int8)
        if t.ndim != 2:
            raise LdpcError("table must be 2-D [iteration][degree]")
        check(load().ldpc_decoder_set_table(self._h, ptr(t, C.c_int8), t.shape[0], t.shape[1]), "set_table")

    def count_errors(self, d_hard: int, frames: int, n_info: int, d_counters: int, d_frame_bits: int = 0,
                     d_frame_ok: int = 0, accumulate: bool = False, stream: int = 0) -> None:
        """Device error/syndrome counters (adds to three uint64 at d_counters)."""
        v = lambda p: C.c_void_p(p) if p else None
        check(load().ldpc_decoder_count_errors(self._h, v(d_hard), frames, n_info, v(d_counters), v(d_frame_bits),
                                               v(d_frame_ok), int(accumulate), v(stream)), "count_errors")

    def simulate(self, ebn0_db: float, n_frames: int, iterations: int, seed: int = 1, snr_idx: int = 0,
                 quant: int = 0) -> dict:
        """One Eb/N0 point on the device: channel -> decode -> count (frames
        k*4W.. of the stream (seed, snr_idx), identical for every decoder)."""
        st = SimStats()
        check(load().ldpc_simulate(self._h, ebn0_db, quant, seed, snr_idx, n_frames, iterations, C.byref(st)),
              "simulate")
        return {"frames": st.frames, "bit_errors": st.bit_errors, "frame_errors": st.frame_errors,
                "syndrome_failures": st.syndrome_failures, "avg_iterations": st.avg_iterations,
                "decode_ms": st.decode_ms, "fer": st.frame_errors / st.frames,
                "ber": st.bit_errors / (st.frames * (self.n - self.k))}

    # ---- state API ----
    def load(self, var_nodes) -> None:
        a = self._in(var_nodes)
        check(load().ldpc_decoder_load(self._h, ptr(a, C.c_int8)), "load")

    def iterate(self, iterations: int = 1) -> None:
        check(load().ldpc_decoder_iterate(self._h, iterations), "iterate")

    def get_state(self):
        post = np.empty(self.n * self.nb_frames, np.uint32)
        msg = np.empty(self.m * self.nb_frames, np.uint32)
        check(load().ldpc_decoder_get_state(self._h, ptr(post, C.c_uint32), ptr(msg, C.c_uint32)), "get_state")
        return post, msg

    def get_hard(self) -> np.ndarray:
        out = np.empty(self.io_bytes, np.uint8)
        check(load().ldpc_decoder_get_hard(self._h, ptr(out, C.c_uint8)), "get_hard")
        return out

    def stop_info(self):
        st = np.empty(self.nb_frames, np.uint8)
        it = np.empty(self.nb_frames, np.int32)
        check(load().ldpc_decoder_get_stop_info(self._h, ptr(st, C.c_uint8), ptr(it, C.c_int)), "stop_info")
        return st, it


class CGPU_Decoder_MS_SIMD(CGPUDecoder):
    """src/decoder_ms/CGPU_Decoder_MS_SIMD.cu"""
    VARIANT = MS


class CGPU_Decoder_OMS_SIMD(CGPUDecoder):
    """src/decoder_oms/CGPU_Decoder_OMS_SIMD.cu"""
    VARIANT = OMS


class CGPU_Decoder_NMS_SIMD(CGPUDecoder):
    """src/decoder_nms/CGPU_Decoder_NMS_SIMD.cu"""
    VARIANT = NMS


class CGPU_Decoder_2NMS_SIMD(CGPUDecoder):
    """src/decoder_2nms/CGPU_Decoder_2NMS_SIMD.cu (early stop on by default)"""
    VARIANT = NMS2


class CGPU_Decoder_MS_SIMD_v2(CGPUDecoder):
    """src/decoder_oms_v2/CGPU_Decoder_MS_SIMD_v2.cu ("xMS"): OMS arithmetic,
    input consumed in the device layout."""
    VARIANT = OMS_V2


class CGPU_Decoder_Learned_OMS(CGPUDecoder):
    VARIANT = LOMS


class CGPU_Decoder_Learned_NMS(CGPUDecoder):
    VARIANT = LNMS


DECODERS = {MS: CGPU_Decoder_MS_SIMD, OMS: CGPU_Decoder_OMS_SIMD, OMS_V2: CGPU_Decoder_MS_SIMD_v2,
            NMS: CGPU_Decoder_NMS_SIMD, NMS2: CGPU_Decoder_2NMS_SIMD, LOMS: CGPU_Decoder_Learned_OMS,
            LNMS: CGPU_Decoder_Learned_NMS}


def make_decoder(code: Code, nb_frames: int, variant: int, **tuning) -> CGPUDecoder:
    return DECODERS[variant](nb_frames, code.n_vars, code.n_checks, code.n_edges, code, **tuning)


# ---------------------------------------------------------------------------
# timer (src/timer/CTimer.cu) and device helpers
# ---------------------------------------------------------------------------
class CTimer:
    def __init__(self, start: bool = False, stream: int = 0):
        self._t = load().ldpc_timer_create(C.c_void_p(stream) if stream else None)
        if not self._t:
            raise LdpcError(_lib.last_error())
        if start:
            self.start()

    def start(self):
        check(load().ldpc_timer_start(self._t), "timer start")

    def stop(self):
        check(load().ldpc_timer_stop(self._t), "timer stop")

    def get_time_ms(self) -> float:
        return float(load().ldpc_timer_elapsed_ms(self._t))

    def __del__(self):
        t = getattr(self, "_t", None)
        if t and _lib._lib is not None:
            _lib._lib.ldpc_timer_destroy(t)
            self._t = None


class DeviceBuffer:
    """Raw device allocation (hipMalloc) owned by Python."""

    def __init__(self, nbytes: int):
        self.nbytes = nbytes
        self.ptr = load().ldpc_device_alloc(nbytes)
        if not self.ptr:
            raise LdpcError(f"device alloc of {nbytes} bytes failed: {_lib.last_error()}")

    def upload(self, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        check(load().ldpc_memcpy_h2d(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes), "h2d")

    def download(self, arr: np.ndarray) -> np.ndarray:
        assert arr.flags.c_contiguous and arr.nbytes <= self.nbytes
        check(load().ldpc_memcpy_d2h(arr.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), arr.nbytes), "d2h")
        return arr

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p and _lib._lib is not None:
            _lib._lib.ldpc_device_free(C.c_void_p(p))
            self.ptr = None


class PinnedBuffer:
    """Page-locked host memory (hipHostMalloc), the counterpart of the
    reference's CTrame buffers (src/trame/CTrame.cpp:38-41).  `array` is a
    uint8 view of it and must not be used after the buffer is dropped."""

    def __init__(self, nbytes: int):
        self.nbytes = nbytes
        self.ptr = load().ldpc_host_alloc_pinned(nbytes)
        if not self.ptr:
            raise LdpcError(f"pinned alloc of {nbytes} bytes failed: {_lib.last_error()}")
        self.array = np.ctypeslib.as_array((C.c_uint8 * nbytes).from_address(self.ptr))

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p and _lib._lib is not None:
            self.array = None
            _lib._lib.ldpc_host_free_pinned(C.c_void_p(p))
            self.ptr = None


class PoolDecoder:
    """Persistent frame-queue decoder: each workgroup keeps 4P frames
    resident, stops every frame on its own (by default when its decisions
    satisfy every check; STOP_V2C_PARITY selects the reference's 2NMS rule;
    else at max_iterations) and refills its slot from a device queue.  Inputs
    and outputs stay on the device (frame-major)."""

    def __init__(self, code: Code, variant: int, max_iterations: int, words_per_block: int = 0,
                 threads: int = 0, table=None, uniform: int = -1, batch: int = 0, stop_rule: int = 0,
                 pairs: int = -1, order: int = ORDER_INPUT):
        """pairs: -1 auto (pairs of frames only when a 4-frame word block does
        not fit LDS), 0 words of 4 frames, 1 pairs of frames.  order: queue
        order of a decode call's frames (ORDER_LEAST_RELIABLE_FIRST starts the
        noisiest frames first; results are the same in either order)."""
        tu = Tuning(words_per_block, threads, -1, {-1: -1, 0: 1, 1: 2}[pairs], uniform, batch)
        h = load().ldpc_pool_create(code.handle, variant, max_iterations, stop_rule, C.byref(tu))
        if not h:
            raise LdpcError(f"ldpc_pool_create failed: {_lib.last_error()}")
        self._h = C.c_void_p(h)
        self.code, self.variant, self.max_iterations, self.stop_rule = code, variant, max_iterations, stop_rule
        if table is not None:
            self.set_table(table)
        if order != ORDER_INPUT:
            self.set_order(order)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.ldpc_pool_destroy(h)
            self._h = None

    def set_table(self, table) -> None:
        t = np.ascontiguousarray(table, dtype=np.int8)
        assert t.ndim == 2
        check(load().ldpc_pool_set_table(self._h, ptr(t, C.c_int8), t.shape[0], t.shape[1]), "pool_set_table")

    def set_order(self, order: int) -> None:
        check(load().ldpc_pool_set_order(self._h, order), "pool_set_order")

    def decode_device(self, d_in: int, frames: int, d_out: int, d_iters: int = 0, workgroups: int = 0,
                      stream: int = 0, graph: bool = False) -> None:
        """graph=True: the step captured in a hipGraph once per buffer set, then replayed."""
        fn = load().ldpc_pool_decode_graph if graph else load().ldpc_pool_decode
        check(fn(self._h, C.c_void_p(d_in), frames, C.c_void_p(d_out), C.c_void_p(d_iters) if d_iters else None,
                 workgroups, C.c_void_p(stream) if stream else None), "pool_decode_graph" if graph else "pool_decode")

    def decode(self, llr: np.ndarray, workgroups: int = 0):
        """Host convenience: frame-major int8 LLRs -> (decisions [F][N], iterations [F])."""
        n = self.code.n_vars
        llr = np.ascontiguousarray(llr, dtype=np.int8).reshape(-1, n)
        F = llr.shape[0]
        d_in, d_out, d_it = DeviceBuffer(llr.nbytes), DeviceBuffer(F * n), DeviceBuffer(4 * F)
        d_in.upload(llr)
        self.decode_device(d_in.ptr, F, d_out.ptr, d_it.ptr, workgroups)
        synchronize()
        hard = d_out.download(np.empty((F, n), np.uint8))
        its = d_it.download(np.empty(F, np.int32))
        return hard, its

    def simulate(self, ebn0_db: float, n_frames: int, seed: int = 1, snr_idx: int = 0, quant: int = 0,
                 batch_frames: int = 0) -> dict:
        """One Eb/N0 point: device channel -> pool decode -> device counting,
        on frames 0..n_frames-1 of the stream (seed, snr_idx)."""
        st = SimStats()
        check(load().ldpc_pool_simulate(self._h, ebn0_db, quant, seed, snr_idx, n_frames, batch_frames,
                                        C.byref(st)), "pool_simulate")
        n_info = self.code.n_vars - self.code.n_checks
        return {"frames": st.frames, "bit_errors": st.bit_errors, "frame_errors": st.frame_errors,
                "syndrome_failures": st.syndrome_failures, "avg_iterations": st.avg_iterations,
                "decode_ms": st.decode_ms, "fer": st.frame_errors / st.frames,
                "ber": st.bit_errors / (st.frames * n_info)}

    @property
    def stream(self) -> int:
        return int(load().ldpc_pool_stream(self._h) or 0)

    def config(self) -> dict:
        P, T, G, U, B, Q = (C.c_int() for _ in range(6))
        check(load().ldpc_pool_config(self._h, C.byref(P), C.byref(T), C.byref(G), C.byref(U), C.byref(B),
                                      C.byref(Q)), "pool_config")
        return {"words_per_block": P.value, "threads": T.value, "max_workgroups": G.value, "uniform_deg": U.value,
                "batch": B.value, "pairs": Q.value}


def pool_autotune(code: Code, variant: int, max_iterations: int, ebn0_db: float = 2.5,
                  sample_frames: int = 32768, stop_rule: int = 0, quant: int = 0) -> dict:
    """Fastest frame-queue tuning for this code (words per workgroup, workgroup
    size, single-degree path and its items in flight), timed on the GPU."""
    tu, ms, n = Tuning(), C.c_double(), C.c_int()
    check(load().ldpc_pool_autotune(code.handle, variant, max_iterations, stop_rule, ebn0_db, quant, sample_frames,
                                    C.byref(tu), C.byref(ms), C.byref(n)), "pool_autotune")
    return {"words_per_block": tu.words_per_block, "threads": tu.threads, "uniform": tu.uniform,
            "batch": tu.batch, "pairs": 1 if tu.post_in_lds == 2 else 0, "ms": ms.value, "tried": n.value}


class MixedDecoder:
    """One decode step over frames of several codes (one frame-queue decoder
    per code, side by side on the GPU, optionally replayed from a hipGraph)."""

    def __init__(self, codes, variant: int, max_iterations: int, tunings=None, max_frames: int = 0,
                 stop_rule: int = 0):
        self.codes = list(codes)
        hs = (C.c_void_p * len(self.codes))(*[c.handle.value for c in self.codes])
        tarr = None
        if tunings is not None:
            tarr = (Tuning * len(self.codes))()
            for i, t in enumerate(tunings):
                t = t or {}
                tarr[i] = Tuning(t.get("words_per_block", 0), t.get("threads", 0), -1,
                                 {-1: -1, 0: 1, 1: 2}[t.get("pairs", -1)], t.get("uniform", -1), t.get("batch", 0))
        h = load().ldpc_mixed_create(hs, len(self.codes), variant, max_iterations, stop_rule, tarr, max_frames)
        if not h:
            raise LdpcError(f"ldpc_mixed_create failed: {_lib.last_error()}")
        self._h = C.c_void_p(h)
        self.variant, self.max_iterations = variant, max_iterations

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.ldpc_mixed_destroy(h)
            self._h = None

    def set_table(self, code_index: int, table) -> None:
        t = np.ascontiguousarray(table, dtype=np.int8)
        assert t.ndim == 2
        check(load().ldpc_mixed_set_table(self._h, code_index, ptr(t, C.c_int8), t.shape[0], t.shape[1]),
              "mixed_set_table")

    def decode_device(self, d_in: int, d_offsets: int, d_code_ids: int, frames: int, d_out: int, d_iters: int = 0,
                      graph: bool = False, stream: int = 0) -> None:
        v = lambda p: C.c_void_p(p) if p else None
        check(load().ldpc_mixed_decode(self._h, v(d_in), v(d_offsets), v(d_code_ids), frames, v(d_out), v(d_iters),
                                       int(graph), v(stream)), "mixed_decode")

    def decode_device_noalloc(self, d_in: int, d_offsets: int, d_code_ids: int, frames: int, d_out: int,
                              d_iters: int = 0, stream: int = 0) -> None:
        """The same step with its scratch in the library's static device array
        (no allocation; frames * codes <= NOALLOC_ENTRIES; calls overlapping on
        one GPU race)."""
        v = lambda p: C.c_void_p(p) if p else None
        check(load().ldpc_mixed_decode_noalloc(self._h, v(d_in), v(d_offsets), v(d_code_ids), frames, v(d_out),
                                               v(d_iters), v(stream)), "mixed_decode_noalloc")

    NOALLOC_ENTRIES = 1 << 21  # LDPC_MIXED_NOALLOC_ENTRIES

    def config(self, code_index: int) -> dict:
        P, T, G, U, B, Q = (C.c_int() for _ in range(6))
        check(load().ldpc_mixed_config(self._h, code_index, C.byref(P), C.byref(T), C.byref(G), C.byref(U),
                                       C.byref(B), C.byref(Q)), "mixed_config")
        return {"words_per_block": P.value, "threads": T.value, "max_workgroups": G.value, "uniform_deg": U.value,
                "batch": B.value, "pairs": Q.value}


def mixed_layout(codes, code_ids):
    """Byte offsets of frames laid out back to back in code_ids order:
    (offsets uint64 [F], total bytes)."""
    n = np.array([c.n_vars for c in codes], np.uint64)[np.asarray(code_ids, np.int64)]
    off = np.zeros(len(n), np.uint64)
    if len(n) > 1:
        np.cumsum(n[:-1], out=off[1:])
    return off, int(n.sum())


def interleave_device(d_in: int, d_words: int, n_vars: int, words: int, stream: int = 0) -> None:
    """Frame-major int8 frames -> the device word layout (word [v][w], byte k =
    frame 4w+k), the input layout of the OMS_v2 decoder."""
    check(load().ldpc_interleave_device(C.c_void_p(d_in), C.c_void_p(d_words), n_vars, words,
                                        C.c_void_p(stream) if stream else None), "interleave_device")


def channel_sigma(snr_db: float, rate: float, es_n0: bool = False) -> float:
    return float(load().ldpc_channel_sigma(snr_db, rate, int(es_n0)))


def channel_generate(d_out: int, n_vars: int, frames: int, seed: int, snr_idx: int, sigma: float,
                     quant: int = 0, layout: int = 0, first_frame: int = 0, stream: int = 0) -> None:
    """Device AWGN channel into d_out (layout 0: frame-major int8, 1: device words)."""
    check(load().ldpc_channel_generate(C.c_void_p(d_out), layout, n_vars, frames, first_frame, seed, snr_idx,
                                       sigma, quant, C.c_void_p(stream) if stream else None), "channel_generate")


def channel_generate_range(d_out: int, n_vars: int, frames: int, seed: int, lo_db: float, hi_db: float,
                           rate: float, quant: int = 0, first_frame: int = 0, snr_idx: int = 0,
                           stream: int = 0) -> None:
    check(load().ldpc_channel_generate_range(C.c_void_p(d_out), n_vars, frames, first_frame, seed, snr_idx, lo_db,
                                             hi_db, rate, quant, C.c_void_p(stream) if stream else None),
          "channel_generate_range")


LEARN_OFFSET, LEARN_SCALE = 0, 1


class Trainer:
    """Float shadow decoder on the GPU (forward + backward HIP kernels) and
    Adam, for per-iteration / per-degree offset or scale tables."""

    def __init__(self, code: Code, mode: int, iterations: int, n_degs: int, batch_frames: int):
        h = load().ldpc_trainer_create(code.handle, mode, iterations, n_degs, batch_frames)
        if not h:
            raise LdpcError(f"ldpc_trainer_create failed: {_lib.last_error()}")
        self._h = C.c_void_p(h)
        self.code, self.mode, self.T, self.dcols, self.frames = code, mode, iterations, n_degs, batch_frames

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.ldpc_trainer_destroy(h)
            self._h = None

    def loss_grad(self, params, d_llr: int = 0, want_grad: bool = True, d_post_out: int = 0):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.shape == (self.T, self.dcols)
        loss = C.c_float()
        grad = np.zeros_like(p) if want_grad else None
        check(load().ldpc_trainer_loss_grad(self._h, ptr(p, C.c_float), C.c_void_p(d_llr) if d_llr else None,
                                            C.byref(loss), ptr(grad, C.c_float) if want_grad else None,
                                            C.c_void_p(d_post_out) if d_post_out else None), "trainer_loss_grad")
        return float(loss.value), grad

    def fit(self, params, steps: int, lr: float = 1e-2, ebn0=(1.5, 2.5), seed: int = 1, first_frame: int = 0,
            quant: int = 0, pmin: float = 0.0, pmax: float = 8.0):
        p = np.ascontiguousarray(params, dtype=np.float32).copy()
        hist = np.zeros(steps, np.float32)
        cfg = LearnConfig(steps, lr, 0.9, 0.999, 1e-8, ebn0[0], ebn0[1], 0.0, seed, first_frame, quant, pmin, pmax)
        check(load().ldpc_trainer_fit(self._h, C.byref(cfg), ptr(p, C.c_float), ptr(hist, C.c_float)),
              "trainer_fit")
        return p, hist


def quantize_offsets(beta: np.ndarray) -> np.ndarray:
    """Float offsets (LSB) -> int8 steps for LDPC_LOMS."""
    return np.clip(np.rint(beta), 0, 31).astype(np.int8)


def quantize_scales(alpha: np.ndarray) -> np.ndarray:
    """Float scales -> numerators over 8 for LDPC_LNMS."""
    return np.clip(np.rint(8 * alpha), 0, 8).astype(np.int8)


def device_count() -> int:
    return int(load().ldpc_device_count())


def synchronize() -> None:
    check(load().ldpc_device_synchronize(), "synchronize")
