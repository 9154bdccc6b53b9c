"""The C-ABI libraries load and export every symbol their headers declare; layouts match the
ctypes mirrors; argument errors come back as codes. No compute without a GPU. CPU only."""
import ctypes as C
import os
import re

import pytest

from specrel_amd import _ffi as F
from specrel_amd import renderer as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    with open(os.path.join(ROOT, "include", header)) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(specrel_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,loader", [("specrel.h", R.hip_lib), ("specrel_scene.h", F.scene_lib),
                                           ("specrel_file.h", F.scene_lib)])
def test_library_exports_every_declared_symbol(header, loader):
    lib = loader()
    names = declared(header)
    assert len(names) > 3
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_mathcheck_library_exports_its_entry():
    lib = C.CDLL(os.path.join(F.LIB_DIR, "libspecrel_mathcheck.so"))
    assert hasattr(lib, "specrel_mathcheck")


def test_python_signatures_cover_the_header():
    assert set(declared("specrel.h")) == set(R.SIGNATURES)


def test_layouts_match_the_header():
    assert C.sizeof(F.Frame) == 800 and F.Frame.haze_origin.offset == 144
    assert F.Attr.misc.offset == 160 and F.Attr.decided_at.offset == 176
    assert C.sizeof(F.SceneArrays) == 88 and F.SceneArrays.light_pos.offset == 72
    assert C.sizeof(F.Stats) == 128
    assert C.sizeof(R.Timings) == 56 and R.Timings.trace_block_waves.offset == 48
    assert R.Timings.trace_variant.offset == 52
    assert C.sizeof(R.ViewOut) == 40 and R.ViewOut.len.offset == 32
    assert C.sizeof(R.Tuning) == 56 and R.Tuning.trace_waves.offset == 40 and R.Tuning.fused_shade.offset == 52
    assert C.sizeof(R.AdjObject) == 24 and C.sizeof(R.AdjointOut) == 216 and C.sizeof(R.Range) == 8
    assert C.sizeof(F.Object) == 32


def test_argument_errors_are_codes_not_aborts():
    lib = R.hip_lib()
    assert lib.specrel_render(None, 0, None, None, None, None, None, 0) == R.EINVAL
    assert lib.specrel_upload_scene(None, 0, None) == R.EINVAL
    assert lib.specrel_sync(None) == R.EINVAL
    assert lib.specrel_set_option(None, 1, 1) == R.EINVAL
    assert lib.specrel_render_batch(None, 0, None, 1, None) == R.EINVAL
    assert lib.specrel_render_batch_resident(None, 0, None, 1, 0) == R.EINVAL
    assert lib.specrel_download_view(None, 0, None, None, None, None, 0) == R.EINVAL
    assert lib.specrel_download(None, None, None, None, None, 0) == R.EINVAL
    assert lib.specrel_autotune(None, 0, None, 1, 1, None) == R.EINVAL
    assert lib.specrel_adjoint(None, None, None, None, 0, None, 0, None, None) == R.EINVAL
    assert lib.specrel_adjoint_batch(None, None, None, None, 0, None, 0, None, None) == R.EINVAL
    assert lib.specrel_patch_scene(None, 0, None, None, 0, None, 0, None) == R.EINVAL
    assert lib.specrel_download_linear(None, 0, None, 0) == R.EINVAL
    assert lib.specrel_upload_scene_versioned(None, 0, None, None) == R.EINVAL
    assert lib.specrel_update_scene(None, 0, None, None, None, 0, None, 0, None) == R.EINVAL
    assert lib.specrel_last_wave_clock(None, None, 0, None, None) == R.EINVAL
    assert C.sizeof(R.WaveClock) == 32 and R.WAVE_CLOCK_DTYPE.itemsize == 32
    assert lib.specrel_render_uniforms(None, 0, None, None, 0) == R.EINVAL
    # GpuUniforms (gpu.rs:120-144): 768 B, fields where the reference's #[repr(C)] puts them
    assert C.sizeof(R.GpuUniforms) == 768
    for name, off in (("light_pos", 80), ("num_nodes", 124), ("frontier_origin", 128), ("num_stars", 164),
                      ("haze_origin", 176), ("flash_at", 256), ("flash_col_width", 512)):
        assert getattr(R.GpuUniforms, name).offset == off, name


def test_create_reports_missing_device():
    lib = R.hip_lib()
    h = C.c_void_p()
    rc = lib.specrel_create(0, C.byref(h))
    if rc == R.OK:  # a gfx950 device is present: nothing more to check here
        lib.specrel_destroy(h)
        return
    assert rc == R.ENODEV
    assert lib.specrel_last_error(None)
    with pytest.raises(RuntimeError):
        R.GpuRenderer([])
