"""The FP32 math kernels of the step core (UCRT-style sinf, cosf, atan2f, powf, acosf in double with explicit FMA).

Three independent pins:
  * the lookup tables are regenerated from the mathematics and must equal the committed include file;
  * the host twin's kernels are compared bit for bit with the reference build's (oracle/_ref) on sampled inputs
    plus the boundary bands of each kernel's argument reduction;
  * the kernels are compared with numpy's double-precision functions: every result is within 1 ulp of the exact
    value rounded to float (a faithful rounding), and exact at the special values the C standard fixes.
"""
import ctypes
import io
import os
import runpy
import sys
from contextlib import redirect_stdout

import numpy as np
import pytest

from conftest import REPO, oracle_or_skip, twin

SIN, COS, ATAN2, POW, ACOS, SAFE_ATAN2 = 0, 1, 2, 3, 4, 5
SIN_LIMIT = np.uint32(0x4B800456).view(np.float32)  # from here on sinf takes the large-argument (libm) path
COS_LIMIT = np.uint32(0x4F490FDB).view(np.float32)


def call(lib_fn, fn, x, y=None):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(x if y is None else y, np.float32)
    out = np.empty_like(x)
    assert lib_fn(fn, x.ctypes.data, y.ctypes.data, out.ctypes.data, x.size) == 0
    return out


def samples(fn, n=200_000, seed=0):
    rng = np.random.default_rng(seed + fn)
    if fn in (SIN, COS):
        bits = rng.integers(0, 0x4B800456, n, dtype=np.uint32)
        x = bits.view(np.float32) * rng.choice(np.array([-1, 1], np.float32), n)
        # The whole finite range beyond the UCRT limits (the large-argument path), and bands around the limits.
        big = rng.integers(0x4B800456, 0x7F800000, n // 2, dtype=np.uint32).view(np.float32)
        big = big * rng.choice(np.array([-1, 1], np.float32), big.size)
        bands = np.concatenate([(np.uint32(0x4B800456) + np.arange(-64, 64, dtype=np.int64)).astype(np.uint32),
                                (np.uint32(0x4F490FDB) + np.arange(-64, 64, dtype=np.int64)).astype(np.uint32)])
        x = np.concatenate([x, big, bands.view(np.float32), -bands.view(np.float32),
                            np.float32(2.0) ** np.arange(24, 128, dtype=np.float32)])
        edges = np.float32(np.pi / 4) * np.arange(-400, 401, dtype=np.float32)
        near = np.concatenate([np.nextafter(edges, np.float32(np.inf)), np.nextafter(edges, np.float32(-np.inf)),
                               edges])
        small = np.array([0.0, -0.0, 1e-30, 2**-14, 2**-7, 2**-7 * 1.0000001, 0.785398, 0.7853982, 0.78539824],
                         np.float32)
        return np.concatenate([x, near, small]), None
    if fn in (ATAN2, SAFE_ATAN2):
        y = (rng.standard_normal(n) * np.exp(rng.uniform(-20, 20, n))).astype(np.float32)
        x = (rng.standard_normal(n) * np.exp(rng.uniform(-20, 20, n))).astype(np.float32)
        sp = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, 3.0, 1e-38], np.float32)
        yy, xx = np.meshgrid(sp, sp)
        return np.concatenate([y, yy.ravel()]), np.concatenate([x, xx.ravel()])
    if fn == POW:
        x = np.abs(rng.standard_normal(n) * np.exp(rng.uniform(-8, 8, n))).astype(np.float32)
        y = rng.uniform(-8, 8, n).astype(np.float32)
        # The step's own call: (1 - drag)^(dt) for drag in [0, 1), dt around 1/120.
        d = rng.uniform(0, 1, 5000).astype(np.float32)
        t = np.full(5000, 1 / 120, np.float32)
        sp = np.array([1.0, 0.0, 2.0, 0.5], np.float32)
        return np.concatenate([x, 1 - d, sp]), np.concatenate([y, t, np.array([0.3, 2.0, -1.0, 0.5], np.float32)])
    if fn == ACOS:
        x = rng.uniform(-1, 1, n).astype(np.float32)
        sp = np.array([1.0, -1.0, 0.0, -0.0, 0.5, -0.5, np.nextafter(np.float32(1), np.float32(0))], np.float32)
        return np.concatenate([x, sp]), None
    raise ValueError(fn)


def exact(fn, x, y):
    xd = x.astype(np.float64)
    if fn == SIN:
        return np.sin(xd)
    if fn == COS:
        return np.cos(xd)
    if fn in (ATAN2, SAFE_ATAN2):
        return np.arctan2(xd, y.astype(np.float64))
    if fn == POW:
        return np.power(xd, y.astype(np.float64))
    return np.arccos(xd)


def test_tables_regenerate():
    gen = os.path.join(REPO, "rocketgpu_amd", "csrc", "core", "gen_math_tables.py")
    buf = io.StringIO()
    with redirect_stdout(buf):
        runpy.run_path(gen, run_name="__main__")
    with open(os.path.join(REPO, "rocketgpu_amd", "csrc", "core", "rg_math_tables.inc")) as fh:
        assert buf.getvalue() == fh.read()


@pytest.mark.parametrize("fn", [SIN, COS, ATAN2, POW, ACOS, SAFE_ATAN2])
def test_twin_math_matches_reference_bits(fn):
    T = twin()
    R = oracle_or_skip()
    x, y = samples(fn)
    a = call(T.rgt_math, fn, x, y)
    b = call(R.lib().rgref_math, fn, x, y)
    same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
    bad = np.flatnonzero(~same)
    assert bad.size == 0, f"{bad.size} differ, e.g. x={x[bad[:4]]} y={None if y is None else y[bad[:4]]}"


@pytest.mark.parametrize("fn", [SIN, COS, ATAN2, POW, ACOS, SAFE_ATAN2])
def test_math_faithful_against_numpy(fn):
    T = twin()
    x, y = samples(fn, n=100_000, seed=100)
    got = call(T.rgt_math, fn, x, y).astype(np.float64)
    want = exact(fn, x, y)
    ok = np.isfinite(want) & (np.abs(want) < 3e38) & (np.abs(want) > 1e-37)
    lo = np.nextafter(want[ok].astype(np.float32), np.float32(-np.inf)).astype(np.float64)
    hi = np.nextafter(want[ok].astype(np.float32), np.float32(np.inf)).astype(np.float64)
    g = got[ok]
    inside = (g >= np.minimum(lo, hi)) & (g <= np.maximum(lo, hi))
    bad = np.flatnonzero(~inside)
    assert bad.size == 0, f"{bad.size} beyond 1 ulp, e.g. x={x[ok][bad[:3]]} got={g[bad[:3]]} want={want[ok][bad[:3]]}"


def test_math_special_values():
    T = twin()
    f = lambda fn, a, b=None: float(call(T.rgt_math, fn, np.array([a], np.float32),
                                         None if b is None else np.array([b], np.float32))[0])
    assert f(SIN, 0.0) == 0.0 and np.signbit(f(SIN, -0.0))
    assert f(COS, 0.0) == 1.0
    assert f(ATAN2, 0.0, 1.0) == 0.0
    assert f(ATAN2, 1.0, 0.0) == np.float32(np.pi / 2)
    assert f(ATAN2, 0.0, -1.0) == np.float32(np.pi)
    assert f(ATAN2, -1.0, 0.0) == -np.float32(np.pi / 2)
    assert f(POW, 2.0, 10.0) == 1024.0 and f(POW, 7.0, 0.0) == 1.0 and f(POW, 1.0, 5.5) == 1.0
    assert f(POW, 4.0, 0.5) == 2.0
    assert f(ACOS, 1.0) == 0.0 and f(ACOS, -1.0) == np.float32(np.pi) and f(ACOS, 0.0) == np.float32(np.pi / 2)


def test_large_arguments_reduce_exactly():
    """Beyond the UCRT limits sinf and cosf reduce exactly: at x = 2^k the results equal the correctly rounded
    sin / cos of the exact float (numpy's double sin has a correct reduction too), to within an ulp, and
    sin(-x) = -sin(x), cos(-x) = cos(x) bit for bit."""
    T = twin()
    x = np.concatenate([np.float32(2.0) ** np.arange(24, 128, dtype=np.float32),
                        np.uint32(0x7F7FFFFF).view(np.float32).reshape(1)])
    s, c = call(T.rgt_math, SIN, x), call(T.rgt_math, COS, x)
    assert np.array_equal(call(T.rgt_math, SIN, -x).view(np.uint32), (-s).view(np.uint32))
    assert np.array_equal(call(T.rgt_math, COS, -x).view(np.uint32), c.view(np.uint32))
    for got, want in ((s, np.sin(x.astype(np.float64))), (c, np.cos(x.astype(np.float64)))):
        w = want.astype(np.float32)
        ulps = np.abs(got.view(np.int32).astype(np.int64) - w.view(np.int32).astype(np.int64))
        assert ulps.max() <= 1
    nan = call(T.rgt_math, SIN, np.array([np.inf, -np.inf, np.nan], np.float32))
    assert np.isnan(nan).all()
