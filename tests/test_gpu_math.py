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


def sky_directions():
    """Uniform directions, and directions whose polar / azimuthal angle puts the sky grid's
    fractional parts (theta / pi * 6, phi / pi * 12) just below, at and just above 0, 0.025 and 1,
    where the device's f32 decision of the grid lines hands over to the f64 routines."""
    rng = np.random.default_rng(12)
    v = rng.normal(size=(1 << 17, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    dirs = [v]
    offs = np.concatenate([-np.logspace(-7, -2, 40), [0.0], np.logspace(-7, -2, 40)])
    for base in (0.0, 0.025, 1.0):
        fr = base + offs
        # azimuth near the thresholds, polar angle spread
        k = rng.integers(-12, 12, fr.size * 40)
        phi = (np.repeat(fr, 40) + k) * np.pi / 12.0
        theta = rng.uniform(0.05, np.pi - 0.05, phi.size)
        dirs.append(np.stack([np.sin(theta) * np.cos(phi), np.cos(theta), np.sin(theta) * np.sin(phi)], 1))
        # polar angle near the thresholds, azimuth spread
        k = rng.integers(0, 6, fr.size * 40)
        theta = np.clip((np.repeat(fr, 40) + k) * np.pi / 6.0, 0.0, np.pi)
        phi = rng.uniform(-np.pi, np.pi, theta.size)
        dirs.append(np.stack([np.sin(theta) * np.cos(phi), np.cos(theta), np.sin(theta) * np.sin(phi)], 1))
    s = np.concatenate(dirs).astype(np.float32)
    d = np.empty((s.shape[0], 4), np.float32)
    d[:, 0] = -1.0  # past light cone: the time component does not enter the sky
    d[:, 1:] = s
    return d


def test_sky_grid_lines(mathcheck):
    """shade_sky on the device (grid lines decided by the f32 acosf / atan2f away from the
    thresholds, by the f64-rounded routines near them) against the oracle's shade_sky, bit for bit."""
    lib = mathcheck
    lib.specrel_skycheck.restype = C.c_int
    lib.specrel_skycheck.argtypes = [FP, FP, C.c_uint64]
    d = np.ascontiguousarray(sky_directions())
    got = np.empty((d.shape[0], 3), np.float32)
    assert lib.specrel_skycheck(d.ctypes.data_as(FP), got.ctypes.data_as(FP), d.shape[0]) == 0
    ref = O.lib()
    want = np.empty_like(got)
    out = np.empty(3, np.float32)
    for i in range(d.shape[0]):
        ref.oracle_shade_sky(d[i].ctypes.data_as(FP), out.ctypes.data_as(FP))
        want[i] = out
    bad = np.flatnonzero((got.view(np.uint32) != want.view(np.uint32)).any(axis=1))
    assert bad.size == 0, f"{bad.size} of {d.shape[0]} sky colours differ, e.g. direction {d[bad[0]]}"
    s = d[:, 1:].astype(np.float64)
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    fu = np.mod(np.arctan2(s[:, 2], s[:, 0]) / np.pi * 12.0, 1.0)
    fv = np.mod(np.arccos(np.clip(s[:, 1], -1.0, 1.0)) / np.pi * 6.0, 1.0)
    near = lambda f: (np.minimum(f, 1.0 - f) < 2e-3) | (np.abs(f - 0.025) < 2e-3)
    print(f"sky: {d.shape[0]} directions identical, {int((near(fu) | near(fv)).sum())} of them within 2e-3 "
          f"of a grid threshold")
