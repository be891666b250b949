"""The device max and min the solves use (maxq / minq: one v_max_f32 / v_min_f32 each) against the
oracle's maxf / minf (src/internal.h:28-45) on every pair of special values (signed zeros, quiet NaNs
with payloads and signs, infinities, denormals, the largest finite values) and on random bit
patterns, a quarter of them equal pairs and an eighth opposite-signed twins. Bit for bit, except where
an operand is a signalling NaN, which the instructions return quieted; the solves apply them only to
results of float arithmetic, which are never signalling."""
import numpy as np
import pytest

import refworld
from twofold_amd import solver

pytestmark = pytest.mark.gpu

SPECIAL = np.array([0x00000000, 0x80000000, 0x3F800000, 0xBF800000, 0x7FC00000, 0xFFC00000, 0x7FC12345,
                    0xFFC54321, 0x7F800000, 0xFF800000, 0x00000001, 0x80000001, 0x007FFFFF, 0x807FFFFF,
                    0x7F7FFFFF, 0xFF7FFFFF, 0x00800000, 0x80800000, 0x3F800001, 0x40000000], np.uint32)
SIGNALLING = np.array([0x7F800001, 0xFFA00000], np.uint32)


def pairs(seed=5, n=4096):
    a = np.repeat(SPECIAL, len(SPECIAL))
    b = np.tile(SPECIAL, len(SPECIAL))
    rng = np.random.default_rng(seed)
    x = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    y = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    y[0::4] = x[0::4]
    y[1::8] = x[1::8] ^ np.uint32(0x80000000)
    x[signalling(x)] |= np.uint32(0x400000)
    y[signalling(y)] |= np.uint32(0x400000)
    return np.concatenate([a, x]), np.concatenate([b, y])


def signalling(u):
    return ((u & 0x7F800000) == 0x7F800000) & ((u & 0x7FFFFF) != 0) & ((u & 0x400000) == 0)


def oracle(a, b):
    lib = refworld._load_oracle()
    fa, fb = a.view(np.float32), b.view(np.float32)
    mx = np.array([lib.tfo_maxf(float(x), float(y)) for x, y in zip(fa, fb)], np.float32)
    mn = np.array([lib.tfo_minf(float(x), float(y)) for x, y in zip(fa, fb)], np.float32)
    return mx.view(np.uint32), mn.view(np.uint32)


def test_device_minmax_is_maxf_minf():
    a, b = pairs()
    assert not signalling(a).any() and not signalling(b).any()
    print(f"{len(a)} pairs", flush=True)
    mx, mn = solver.device_minmax(a.view(np.float32), b.view(np.float32))
    want_mx, want_mn = oracle(a, b)
    assert np.array_equal(mx.view(np.uint32), want_mx), np.flatnonzero(mx.view(np.uint32) != want_mx)[:8]
    assert np.array_equal(mn.view(np.uint32), want_mn), np.flatnonzero(mn.view(np.uint32) != want_mn)[:8]


def test_signalling_nan_comes_back_quiet():
    """The one documented difference: a signalling NaN operand gives a quiet NaN where maxf / minf
    give the other operand."""
    a = np.concatenate([SIGNALLING, np.full(2, 0x3F800000, np.uint32)])
    b = np.concatenate([np.full(2, 0x3F800000, np.uint32), SIGNALLING])
    mx, mn = solver.device_minmax(a.view(np.float32), b.view(np.float32))
    assert np.isnan(mx).all() and np.isnan(mn).all()
    assert not signalling(mx.view(np.uint32)).any()
