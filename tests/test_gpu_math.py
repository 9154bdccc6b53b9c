"""The device's f32 transcendental helpers (exp, sin, cos, atan2, acos: the device math library's
double routine rounded once, as the kernels call them) against the oracle's definitions of the
same helpers (glibc's double routine rounded once), bit for bit, on dense sweeps of at least 2^24
points per function over the ranges the shading uses: fog and star exponents, hand angles, haze
ripples, dial and sky directions. The image parity of every other GPU test rests on these agreeing.
"""
import ctypes as C
import os

import numpy as np
import pytest

from specrel_amd import _ffi as F
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

N = 1 << 24
FP = C.POINTER(C.c_float)
FN = {"exp": 0, "sin": 1, "cos": 2, "atan2": 3, "acos": 4}


@pytest.fixture(scope="module")
def mathcheck():
    path = os.path.join(F.LIB_DIR, "libspecrel_mathcheck.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m specrel_amd.build`")
    lib = C.CDLL(path)
    lib.specrel_mathcheck.restype = C.c_int
    lib.specrel_mathcheck.argtypes = [C.c_int, FP, FP, FP, C.c_uint64]
    return lib


def device(lib, fn, a, b=None):
    a = np.ascontiguousarray(a, np.float32)
    b = None if b is None else np.ascontiguousarray(b, np.float32)
    out = np.empty_like(a)
    rc = lib.specrel_mathcheck(FN[fn], a.ctypes.data_as(FP), None if b is None else b.ctypes.data_as(FP),
                               out.ctypes.data_as(FP), a.size)
    assert rc == 0, rc
    return out


def sweep(lo, hi, n=N):
    return np.linspace(lo, hi, n, dtype=np.float64).astype(np.float32)


def uniform(lo, hi, seed, n=N):
    return np.random.default_rng(seed).uniform(lo, hi, n).astype(np.float32)


def compare(lib, fn, a, b=None):
    got = device(lib, fn, a, b)
    want = O.math_vec(fn, a, b)
    same = got.view(np.uint32) == want.view(np.uint32)
    both_nan = np.isnan(got) & np.isnan(want)
    bad = ~(same | both_nan)
    where = np.flatnonzero(bad)[:5]
    assert not bad.any(), (f"{fn}: {int(bad.sum())} of {a.size} differ, e.g. at "
                           + ", ".join(f"{a[i]!r}{'' if b is None else ', ' + repr(b[i])}: {got[i]!r} vs {want[i]!r}"
                                       for i in where))
    return a.size


@pytest.mark.parametrize("lo,hi,seed", [(-100.0, 90.0, None), (-2.0, 0.0, 1), (-88.0, 88.0, 2)])
def test_exp(mathcheck, lo, hi, seed):
    a = sweep(lo, hi) if seed is None else uniform(lo, hi, seed)
    print(f"exp over [{lo}, {hi}]: {compare(mathcheck, 'exp', a)} points identical", flush=True)


@pytest.mark.parametrize("fn", ["sin", "cos"])
@pytest.mark.parametrize("lo,hi,seed", [(-1000.0, 1000.0, None), (-8.0, 8.0, 3)])
def test_sin_cos(mathcheck, fn, lo, hi, seed):
    a = sweep(lo, hi) if seed is None else uniform(lo, hi, seed)
    print(f"{fn} over [{lo}, {hi}]: {compare(mathcheck, fn, a)} points identical", flush=True)


def test_atan2(mathcheck):
    g = sweep(-1.0, 1.0, 4096)
    y, x = np.meshgrid(g, g, indexing="ij")
    n = compare(mathcheck, "atan2", y.ravel(), x.ravel())
    sy = uniform(-0.4, 0.4, 4)  # dial offsets around the pivot
    sx = uniform(-0.4, 0.4, 5)
    n += compare(mathcheck, "atan2", sy, sx)
    edge = np.array([0.0, -0.0, 1.0, -1.0, 1e-30, -1e-30, np.inf, -np.inf], np.float32)
    ey, ex = np.meshgrid(edge, edge, indexing="ij")
    n += compare(mathcheck, "atan2", ey.ravel(), ex.ravel())
    print(f"atan2: {n} points identical")


def test_acos(mathcheck):
    n = compare(mathcheck, "acos", sweep(-1.0, 1.0))
    n += compare(mathcheck, "acos", uniform(-1.0, 1.0, 6))
    print(f"acos: {n} points identical")
