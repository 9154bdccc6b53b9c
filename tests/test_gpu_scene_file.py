"""GPU upload of a scene file (specrel_upload_scene_file): the render of a scene loaded from its
file is bit-identical to the render of the in-memory upload and to the oracle's; bad files are
refused with SPECREL_EINVAL and leave the context usable."""
import numpy as np
import pytest

from specrel_amd import _ffi as F
from specrel_amd import scene as S
from specrel_amd.renderer import SpecrelError
from oracle import pyoracle as O
from tests.helpers import assert_same, frame_for

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,params", [("twin", None), ("large", dict(n_objects=6, slices=16, n_stars=300))])
def test_file_upload_renders_like_the_in_memory_scene(gpu_ctx, tmp_path, name, params):
    p = S.gen_params(name, **params) if params else None
    sc = S.Scene.generate(name, p)
    path = tmp_path / f"{name}.srs"
    sc.save(path)
    fr = frame_for(name, 160, 100, params=p)
    gpu_ctx.upload(0, sc)
    mem = gpu_ctx.render_frame(0, fr, want_hits=True)
    gpu_ctx.upload_file(1, path)
    got = gpu_ctx.render_frame(1, fr, want_hits=True)
    for k in ("rgb", "hit", "lam", "tau"):
        a, b = mem[k], got[k]
        same = (a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype == np.float32 else a == b
        assert same.all(), k
    f = S.SceneFile(path)
    assert_same(got, O.render(f, fr), f"{name} from file vs oracle")
    f.close()


def test_bad_file_is_refused_and_the_context_stays_usable(gpu_ctx, tmp_path):
    sc = S.Scene.generate("twin")
    path = tmp_path / "twin.srs"
    sc.save(path)
    raw = bytearray(path.read_bytes())
    geoms = F.FileSection.from_buffer_copy(bytes(raw[128 + 32:128 + 64]))
    assert geoms.kind == 2
    raw[geoms.offset + 1000] ^= 1  # inside a record: only the checksum sees it
    bad = tmp_path / "bad.srs"
    bad.write_bytes(bytes(raw))
    with pytest.raises(SpecrelError, match="checksum mismatch"):
        gpu_ctx.upload_file(0, bad)
    with pytest.raises(SpecrelError, match="cannot open"):
        gpu_ctx.upload_file(0, tmp_path / "absent.srs")
    gpu_ctx.upload_file(0, path)
    fr = frame_for("twin", 64, 40)
    assert_same(gpu_ctx.render_frame(0, fr, want_hits=True), O.render(sc, fr), "after refusals")
