"""Pins of the oracle's per-byte operations.

The reference's ops are NVIDIA inline-PTX intrinsics documented by the table in
src/decoder_ms/cuda/simd_functions.h:89-131.  Each op is checked exhaustively
over all 65 536 byte pairs (in every lane) against numpy int16 arithmetic of
that documented definition, plus the special values the decoder depends on
(wrap-around |-128|, unsigned order of 0x80, strict > 0)."""
import numpy as np
import pytest

from oracle import oracle as O

rng = np.random.default_rng(7)
A, B = np.meshgrid(np.arange(256, dtype=np.int64), np.arange(256, dtype=np.int64), indexing="ij")
A, B = A.ravel(), B.ravel()


def s8(x):
    return ((x + 128) % 256) - 128


DEFS = {
    "vaddss4": lambda a, b: np.clip(s8(a) + s8(b), -128, 127) & 0xFF,
    "vsubss4": lambda a, b: np.clip(s8(a) - s8(b), -128, 127) & 0xFF,
    "vsubus4": lambda a, b: np.clip(a - b, 0, 255),
    "vminu4": lambda a, b: np.minimum(a, b),
    "vmaxu4": lambda a, b: np.maximum(a, b),
    "vcmpgts4": lambda a, b: np.where(s8(a) > s8(b), 0xFF, 0),
    "vcmpeq4": lambda a, b: np.where(a == b, 0xFF, 0),
    "vcmpne4": lambda a, b: np.where(a != b, 0xFF, 0),
    "vsetgts4": lambda a, b: np.where(s8(a) > s8(b), 1, 0),
}


def run_binary(name, a, b):
    f = getattr(O.lib(), "orc_" + name)
    return np.array([f(int(x), int(y)) for x, y in zip(a, b)], dtype=np.int64)


@pytest.mark.parametrize("name", sorted(DEFS))
def test_binary_ops_exhaustive(name):
    # all byte pairs; lane k of word i holds pair (i*4+k) % 65536 so every lane sees every pair
    idx = np.arange(65536)
    words_a = np.zeros(65536, np.int64)
    words_b = np.zeros(65536, np.int64)
    exp = np.zeros(65536, np.int64)
    for k in range(4):
        j = (idx * 4 + k * 16411) % 65536
        words_a |= A[j] << (8 * k)
        words_b |= B[j] << (8 * k)
        exp |= DEFS[name](A[j], B[j]) << (8 * k)
    got = run_binary(name, words_a, words_b)
    np.testing.assert_array_equal(got, exp)


def test_unary_ops_exhaustive():
    x = np.arange(256, dtype=np.int64)
    words = x | (((x + 1) % 256) << 8) | (((x + 77) % 256) << 16) | (((x + 200) % 256) << 24)
    for name, f in (("vabs4", lambda v: np.abs(s8(v)) & 0xFF), ("vneg4", lambda v: (-s8(v)) & 0xFF)):
        got = np.array([getattr(O.lib(), "orc_" + name)(int(w)) for w in words], dtype=np.int64)
        exp = sum(f((x + o) % 256) << (8 * k) for k, o in enumerate((0, 1, 77, 200)))
        np.testing.assert_array_equal(got, exp)


def test_special_values():
    L = O.lib()
    assert L.orc_vabs4(0x80808080) == 0x80808080          # |-128| wraps to 0x80 (=128 unsigned)
    assert L.orc_vminu4(0x80808080, 0x7F7F7F7F) == 0x7F7F7F7F  # 128 > 127 unsigned: never a new min
    assert L.orc_vsubss4(0x80, 0x01) == 0x80                # -128 - 1 saturates
    assert L.orc_vaddss4(0x7F, 0x01) == 0x7F                # 127 + 1 saturates
    assert L.orc_vcmpgts4(0x00, 0x00) == 0x00               # zero is not positive
    assert L.orc_vneg4(0x80) == 0x80                        # -(-128) wraps
    assert L.orc_vsubus4(0x00010203, 0x01010101) == 0x00000102
    # lanes are independent: a carry must not cross bytes
    assert L.orc_vaddss4(0x007F7F00, 0x00010100) == 0x007F7F00
