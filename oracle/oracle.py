"""ctypes face of the CPU oracle (oracle/ldpc_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libldpc_oracle.so"
MS, OMS, OMS_V2, NMS, NMS2, LOMS, LNMS = range(7)

_i_p = C.POINTER(C.c_int)
_u32_p = C.POINTER(C.c_uint32)
_u8_p = C.POINTER(C.c_uint8)
_i8_p = C.POINTER(C.c_int8)
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            from . import build
            build.build()
        L = C.CDLL(str(LIB_PATH))
        for op in ("vaddss4", "vsubss4", "vsubus4", "vminu4", "vmaxu4", "vcmpgts4", "vcmpeq4",
                   "vcmpne4", "vsetgts4"):
            f = getattr(L, "orc_" + op)
            f.restype, f.argtypes = C.c_uint32, [C.c_uint32, C.c_uint32]
        for op in ("vabs4", "vneg4"):
            f = getattr(L, "orc_" + op)
            f.restype, f.argtypes = C.c_uint32, [C.c_uint32]
        dec_args = [C.c_int, C.c_int, _i_p, _u32_p, C.c_int, C.c_size_t, _u32_p, _u32_p, C.c_int, C.c_int,
                    C.c_int, C.c_int, _u8_p, _i8_p, C.c_int, C.c_int, C.c_int, _i_p]
        L.orc_decode.restype, L.orc_decode.argtypes = C.c_int, dec_args
        L.orc_decode_mt.restype, L.orc_decode_mt.argtypes = C.c_int, dec_args + [C.c_int]
        L.orc_interleave.restype, L.orc_interleave.argtypes = None, [_i8_p, _u32_p, C.c_int, C.c_size_t]
        L.orc_deinterleave_hd.restype, L.orc_deinterleave_hd.argtypes = None, [_u32_p, _u8_p, C.c_int, C.c_size_t]
        L.orc_philox4x32_10.restype = None
        L.orc_philox4x32_10.argtypes = [_u32_p, _u32_p, _u32_p]
        L.orc_sigma.restype, L.orc_sigma.argtypes = C.c_double, [C.c_double, C.c_double]
        L.orc_gauss.restype, L.orc_gauss.argtypes = C.c_double, [C.c_uint64, C.c_uint32, C.c_uint64]
        L.orc_llr.restype = C.c_int8
        L.orc_llr.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_double, C.c_int]
        L.orc_channel_frames.restype = None
        L.orc_channel_frames.argtypes = [_i8_p, C.c_int, C.c_size_t, C.c_size_t, C.c_uint64, C.c_uint32,
                                         C.c_double, C.c_int]
        L.orc_channel_frames_range.restype = None
        L.orc_channel_frames_range.argtypes = [_i8_p, C.c_int, C.c_size_t, C.c_size_t, C.c_uint64, C.c_uint32,
                                               C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_count_errors.restype = C.c_longlong
        L.orc_count_errors.argtypes = [_u8_p, C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_longlong), _u32_p]
        L.orc_shadow_loss_grad.restype = C.c_double
        L.orc_shadow_loss_grad.argtypes = [C.c_int, C.c_int, _i_p, _u32_p, C.c_int, C.c_int, C.c_int,
                                           C.POINTER(C.c_float), _i8_p, C.c_size_t, C.POINTER(C.c_float)]
        L.orc_shadow_forward.restype = C.c_double
        L.orc_shadow_forward.argtypes = L.orc_shadow_loss_grad.argtypes
        L.orc_decode_frames.restype = C.c_int
        L.orc_decode_frames.argtypes = [C.c_int, C.c_int, _i_p, _u32_p, C.c_int, _i8_p, C.c_size_t, C.c_int, _i8_p,
                                        C.c_int, C.c_int, _u8_p, _i_p, C.c_int, C.c_int]
        L.orc_syndrome.restype = None
        L.orc_syndrome.argtypes = [_u8_p, C.c_int, C.c_int, _i_p, _u32_p, C.c_size_t, _u8_p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


def op(name: str, *args) -> int:
    return int(getattr(lib(), "orc_" + name)(*args))


def interleave(llr: np.ndarray, n: int) -> np.ndarray:
    llr = np.ascontiguousarray(llr, dtype=np.int8).reshape(-1)
    W = llr.size // (4 * n)
    post = np.empty(n * W, np.uint32)
    lib().orc_interleave(_p(llr, C.c_int8), _p(post, C.c_uint32), n, W)
    return post


def deinterleave_hd(post: np.ndarray, n: int) -> np.ndarray:
    post = np.ascontiguousarray(post, dtype=np.uint32)
    W = post.size // n
    out = np.empty(4 * W * n, np.uint8)
    lib().orc_deinterleave_hd(_p(post, C.c_uint32), _p(out, C.c_uint8), n, W)
    return out


STOP_SYNDROME, STOP_V2C_PARITY = 0, 1


def decode_frames(deg, pos, n_vars: int, variant: int, llr: np.ndarray, max_iters: int, table=None,
                  threads: int = 1, stop_rule: int = STOP_SYNDROME):
    """Per-frame stop decode (the frame-queue decoder's model): returns
    (hard [F][N] uint8, iterations [F] int32)."""
    deg = np.ascontiguousarray(deg, dtype=np.int32)
    pos = np.ascontiguousarray(pos, dtype=np.uint32)
    llr = np.ascontiguousarray(llr, dtype=np.int8).reshape(-1, n_vars)
    F = llr.shape[0]
    hard = np.empty((F, n_vars), np.uint8)
    its = np.empty(F, np.int32)
    t = None if table is None else np.ascontiguousarray(table, dtype=np.int8)
    targs = (_p(t, C.c_int8), t.shape[0], t.shape[1]) if t is not None else (None, 0, 0)
    rc = lib().orc_decode_frames(n_vars, deg.size, _p(deg, C.c_int), _p(pos, C.c_uint32), variant,
                                 _p(llr, C.c_int8), F, max_iters, *targs, _p(hard, C.c_uint8), _p(its, C.c_int),
                                 stop_rule, threads)
    if rc != 0:
        raise ValueError("oracle rejected the arguments")
    return hard, its


class OracleDecoder:
    """Serial layered decode of W words on the CPU, state kept in the
    reference layouts: post[v*W + w], msg[e*W + w] (table-order edges)."""

    def __init__(self, deg, pos, n_vars: int, W: int, variant: int, q1_split: int = -1,
                 early_term: int = -1, table=None):
        self.deg = np.ascontiguousarray(deg, dtype=np.int32)
        self.pos = np.ascontiguousarray(pos, dtype=np.uint32)
        self.N, self.C, self.W = n_vars, self.deg.size, W
        self.E = int(self.deg.sum())
        self.variant = variant
        if q1_split < 0:
            i = 1
            while i < self.C and self.deg[i] == self.deg[0]:
                i += 1
            q1_split = i
        self.q1_split = q1_split
        self.early_term = (variant == NMS2) if early_term < 0 else bool(early_term)
        self.table = None if table is None else np.ascontiguousarray(table, dtype=np.int8)
        self.post = np.zeros(self.N * W, np.uint32)
        self.msg = np.zeros(self.E * W, np.uint32)
        self.stopped = np.zeros(W, np.uint8)
        self.iters_run = np.zeros(W, np.int32)
        self.first = True
        self.done = 0

    def load_words(self, post: np.ndarray) -> None:
        self.post[:] = np.ascontiguousarray(post, dtype=np.uint32).reshape(-1)
        self.msg[:] = 0
        self.stopped[:] = 0
        self.iters_run[:] = 0
        self.first = True
        self.done = 0

    def load(self, llr: np.ndarray) -> None:
        self.load_words(interleave(llr, self.N))

    def iterate(self, iters: int, threads: int = 1) -> None:
        ran = np.zeros(self.W, np.int32)
        t = self.table
        targs = (_p(t, C.c_int8), t.shape[0], t.shape[1]) if t is not None else (None, 0, 0)
        args = [self.N, self.C, _p(self.deg, C.c_int), _p(self.pos, C.c_uint32), self.variant, self.W,
                _p(self.post, C.c_uint32), _p(self.msg, C.c_uint32), iters, int(self.first), self.q1_split,
                int(self.early_term), _p(self.stopped, C.c_uint8), *targs, self.done, _p(ran, C.c_int)]
        rc = lib().orc_decode_mt(*args, threads) if threads > 1 else lib().orc_decode(*args)
        if rc != 0:
            raise ValueError("oracle rejected the arguments")
        self.iters_run += ran
        self.first = False
        self.done += iters

    def hard(self) -> np.ndarray:
        return deinterleave_hd(self.post, self.N)


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.empty(4, np.uint32)
    lib().orc_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(out, C.c_uint32))
    return out


def sigma(ebn0_db: float, rate: float) -> float:
    return float(lib().orc_sigma(ebn0_db, rate))


def channel_frames(n: int, f0: int, nf: int, seed: int, snr_idx: int, sig: float, quant: int = 0) -> np.ndarray:
    out = np.empty(nf * n, np.int8)
    lib().orc_channel_frames(_p(out, C.c_int8), n, f0, nf, seed, snr_idx, sig, quant)
    return out


def channel_frames_range(n, f0, nf, seed, snr_idx, lo, hi, rate=0.5, quant=0):
    out = np.empty(nf * n, np.int8)
    lib().orc_channel_frames_range(_p(out, C.c_int8), n, f0, nf, seed, snr_idx, lo, hi, rate, quant)
    return out


def count_errors(hard: np.ndarray, n: int, n_info: int):
    hard = np.ascontiguousarray(hard, dtype=np.uint8)
    frames = hard.size // n
    be = C.c_longlong()
    per = np.empty(frames, np.uint32)
    fe = lib().orc_count_errors(_p(hard, C.c_uint8), n, frames, n_info, C.byref(be), _p(per, C.c_uint32))
    return int(fe), int(be.value), per


def syndrome_ok(hard: np.ndarray, n: int, deg, pos) -> np.ndarray:
    hard = np.ascontiguousarray(hard, dtype=np.uint8)
    deg = np.ascontiguousarray(deg, dtype=np.int32)
    pos = np.ascontiguousarray(pos, dtype=np.uint32)
    frames = hard.size // n
    ok = np.empty(frames, np.uint8)
    lib().orc_syndrome(_p(hard, C.c_uint8), n, deg.size, _p(deg, C.c_int), _p(pos, C.c_uint32), frames,
                       _p(ok, C.c_uint8))
    return ok


def shadow_loss_grad(deg, pos, n, mode, params, llr, want_grad=True):
    """Float shadow decoder loss (and gradient) over frame-major int8 LLRs."""
    deg = np.ascontiguousarray(deg, dtype=np.int32)
    pos = np.ascontiguousarray(pos, dtype=np.uint32)
    prm = np.ascontiguousarray(params, dtype=np.float32)
    llr = np.ascontiguousarray(llr, dtype=np.int8).reshape(-1)
    T, dcols = prm.shape
    grad = np.zeros_like(prm) if want_grad else None
    loss = lib().orc_shadow_loss_grad(n, deg.size, _p(deg, C.c_int), _p(pos, C.c_uint32), mode, T, dcols,
                                      _p(prm, C.c_float), _p(llr, C.c_int8), llr.size // n,
                                      _p(grad, C.c_float) if want_grad else None)
    return float(loss), grad


def shadow_forward(deg, pos, n, mode, params, llr):
    """Float shadow decoder: (mean loss, final posterior [frames][n])."""
    deg = np.ascontiguousarray(deg, dtype=np.int32)
    pos = np.ascontiguousarray(pos, dtype=np.uint32)
    prm = np.ascontiguousarray(params, dtype=np.float32)
    llr = np.ascontiguousarray(llr, dtype=np.int8).reshape(-1)
    F = llr.size // n
    out = np.empty(F * n, np.float32)
    loss = lib().orc_shadow_forward(n, deg.size, _p(deg, C.c_int), _p(pos, C.c_uint32), mode, prm.shape[0],
                                    prm.shape[1], _p(prm, C.c_float), _p(llr, C.c_int8), F, _p(out, C.c_float))
    return float(loss), out.reshape(F, n)
