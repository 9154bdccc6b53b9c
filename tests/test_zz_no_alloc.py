"""specrel_render_noalloc (no counterpart in the reference: for callers that cannot allocate):
through the C ABI, into device buffers the test owns, its images are bit-identical to
specrel_render's for two calls in a row, the second with a smaller view; views that need more
than the fixed scratch area, options that need buffers of their own, and downloads after it are
refused."""
import ctypes as C

import numpy as np
import pytest

from specrel_amd import renderer as R
from specrel_amd import scene as S

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """n 32-bit words of device memory from the HIP runtime the library itself uses (the test's
    own allocation, outside the entry point under test), filled with 0xFFFFFFFF."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            R.hip_lib()  # loads libamdhip64 as a dependency of the renderer
            cls._hip = C.CDLL("libamdhip64.so.7")
            cls._hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            cls._hip.hipFree.argtypes = [C.c_void_p]
            cls._hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
            cls._hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        return cls._hip

    def __init__(self, n):
        self.n = n
        self.p = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(self.p), 4 * n) == 0
        assert self.hip().hipMemset(self.p, 0xFF, 4 * n) == 0

    def data_ptr(self):
        return self.p.value

    def numpy(self, dtype):
        out = np.zeros(self.n, np.uint32)
        assert self.hip().hipMemcpy(out.ctypes.data, self.p, 4 * self.n, 2) == 0  # device to host
        return out.view(dtype)

    def __del__(self):
        if self.p.value:
            self.hip().hipFree(self.p)
            self.p = C.c_void_p()


def device_outputs(n):
    return [DeviceBuffer(n) for _ in range(4)]


def host(t, dtype):
    return t.numpy(dtype)


def test_noalloc_render_matches_render(gpu_ctx):
    p = S.gen_params("large", n_objects=16, slices=16, n_stars=10000)
    sc = S.Scene.generate("large", p)
    gpu_ctx.upload(0, sc)
    cam = S.default_camera("large", p)
    big = S.build_frame(cam, S.RenderConfig(1280, 800))
    cam2 = S.default_camera("large", p)
    cam2.tetrad.boost_local((0.2, -0.4, 0.3))
    small = S.build_frame(cam2, S.RenderConfig(320, 200, S.RelativisticShaded, S.ShadowLorentzInvariant,
                                               S.CameraSimultaneity))
    print("noalloc: 1280x800 then 320x200, 10,000 stars", flush=True)
    outs = []
    for fr in (big, small):  # two calls in a row, the second smaller
        n = fr.size[0] * fr.size[1]
        d = device_outputs(n)
        rc = gpu_ctx.lib.specrel_render_noalloc(gpu_ctx.handle, 0, C.byref(fr), d[0].data_ptr(), d[1].data_ptr(),
                                                d[2].data_ptr(), d[3].data_ptr(), n)
        assert rc == R.OK, gpu_ctx.lib.specrel_last_error(gpu_ctx.handle).decode()
        gpu_ctx.sync()
        outs.append(d)
    for fr, d in zip((big, small), outs):
        w, h = fr.size[0], fr.size[1]
        ref = gpu_ctx.render_frame(0, fr, want_hits=True)
        got = {"rgb": host(d[0], np.uint32), "hit": host(d[1], np.uint32), "lam": host(d[2], np.float32),
               "tau": host(d[3], np.float32)}
        assert (ref["hit"] != 0xFFFFFFFF).sum() > 100
        for k in ("rgb", "hit", "lam", "tau"):
            assert np.array_equal(got[k].view(np.uint32), ref[k].reshape(-1).view(np.uint32)), f"{w}x{h}: {k}"
    # RGB alone, once more at the large size after the small one
    d = device_outputs(1280 * 800)
    gpu_ctx.render_noalloc(0, big, d[0].data_ptr())
    gpu_ctx.sync()
    assert np.array_equal(host(d[0], np.uint32), gpu_ctx.render_frame(0, big)["rgb"].reshape(-1))


def test_noalloc_render_refusals(gpu_ctx):
    sc = S.Scene.generate("twin")
    gpu_ctx.upload(0, sc)
    lib, h = gpu_ctx.lib, gpu_ctx.handle
    fr = S.build_frame(sc.default_camera(), S.RenderConfig(160, 100))
    d = device_outputs(160 * 100)
    ptrs = [t.data_ptr() for t in d]
    assert lib.specrel_render_noalloc(h, 0, C.byref(fr), ptrs[0], None, None, None, 160 * 99) == R.EINVAL
    assert lib.specrel_render_noalloc(h, 0, C.byref(fr), ptrs[0], ptrs[1], None, None, 160 * 100) == R.EINVAL
    huge = S.build_frame(sc.default_camera(), S.RenderConfig(4096, 4096))
    assert lib.specrel_render_noalloc(h, 0, C.byref(huge), ptrs[0], None, None, None, 4096 * 4096) == R.EINVAL
    assert "bytes of scratch" in lib.specrel_last_error(h).decode()
    try:
        gpu_ctx.set_option(R.OPT_STATS, 1)
        assert lib.specrel_render_noalloc(h, 0, C.byref(fr), ptrs[0], None, None, None, 160 * 100) == R.EINVAL
    finally:
        gpu_ctx.set_option(R.OPT_STATS, 0)
    assert lib.specrel_render_noalloc(h, 0, C.byref(fr), ptrs[0], None, None, None, 160 * 100) == R.OK
    gpu_ctx.sync()
    with pytest.raises(R.SpecrelError):
        gpu_ctx.download()
    ref = gpu_ctx.render_frame(0, fr)
    assert np.array_equal(host(d[0], np.uint32), ref["rgb"].reshape(-1))
    assert gpu_ctx.download()["rgb"].shape == (100, 160)  # a context render again: downloads work
