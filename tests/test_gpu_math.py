"""The exact-arithmetic helpers of the tick evaluated on the device (bchip_eval_math) against the
host: exp against glibc ::exp bit for bit (both table sources), the three correctly rounded
division helpers against IEEE division, the Bernoulli function against the oracle's. The
operands span the whole double range of exp, float arguments, small and subnormal ones, and
division operands over wide exponent ranges with subnormal and near-overflow quotients."""
import ctypes as C

import numpy as np
import pytest

from bioneuroncore_amd import backend
import harness as H

pytestmark = pytest.mark.gpu
N = 2_000_000


def _dev(fn, a, b=None):
    lib = backend.load_library()
    a = np.ascontiguousarray(a, np.float64)
    out = np.zeros_like(a)
    bb = None if b is None else np.ascontiguousarray(b, np.float64)
    assert lib.bchip_eval_math(fn, a.ctypes.data, None if bb is None else bb.ctypes.data, out.ctypes.data, len(a)) == 0
    return out


def _host(name, x):
    lib = H.oracle_lib()
    f = getattr(lib, name)
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros_like(x)
    f(x.ctypes.data, y.ctypes.data, len(x))
    return y


def _exp_inputs(rng):
    parts = [rng.uniform(-745.0, 709.7, N // 5),
             rng.uniform(-60.0, 60.0, N // 5).astype(np.float32).astype(np.float64),
             rng.uniform(-1.0, 1.0, N // 5),
             np.ldexp(rng.uniform(-1.0, 1.0, N // 5), -rng.integers(0, 60, N // 5)),
             -rng.uniform(0.0, 1.0, N // 5).astype(np.float32).astype(np.float64)]
    return np.concatenate(parts)


@pytest.mark.parametrize("fn", [0, 1])
def test_device_exp_is_glibc(fn):
    x = _exp_inputs(np.random.default_rng(5 + fn))
    assert np.array_equal(_dev(fn, x).view(np.uint64), _host("oracle_exp_n", x).view(np.uint64))


def _f32_bits(rng, n, emin, emax):
    e = rng.integers(emin, emax, n).astype(np.uint32)
    m = rng.integers(0, 1 << 23, n).astype(np.uint32)
    s = rng.integers(0, 2, n).astype(np.uint32)
    return ((s << 31) | (e << 23) | m).view(np.float32)


def test_device_divisions_are_ieee():
    rng = np.random.default_rng(9)
    # float / float through the double reciprocal: subnormal through huge exponents
    a = _f32_bits(rng, N, 0, 255)
    b = _f32_bits(rng, N, 1, 255)
    want = (a / b).astype(np.float32)
    got = _dev(3, a.astype(np.float64), b.astype(np.float64)).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # double / widened float (div_f24): flux-like numerators, volume-like divisors
    A = np.ldexp(rng.uniform(-1.0, 1.0, N), rng.integers(-300, 300, N))
    B = _f32_bits(rng, N, 1, 254).astype(np.float64)
    assert np.array_equal(_dev(4, A, B).view(np.uint64), (A / B).view(np.uint64))
    # double / double through the correctly rounded reciprocal (div_rcp)
    D = np.ldexp(rng.uniform(0.5, 1.0, N), rng.integers(-500, 500, N)) * np.where(rng.random(N) < 0.5, -1, 1)
    assert np.array_equal(_dev(2, A, D).view(np.uint64), (A / D).view(np.uint64))


def test_device_bernoulli_is_the_oracles():
    rng = np.random.default_rng(13)
    x = np.concatenate([rng.uniform(-60.0, 60.0, N // 2), rng.uniform(-0.01, 0.01, N // 2)])
    x = x.astype(np.float32).astype(np.float64)
    got = _dev(5, x).astype(np.float32)
    want = _host("oracle_bernoulli_n", x).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
