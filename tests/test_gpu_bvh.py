"""Device BVH build (specrel_build_bvh) against the host build (specrel_build_bvh_host, the
scene library's Bvh4D::build): nodes and leaf order byte for byte, on generated scenes, on the
records of a device extrusion, on synthetic bounds with clustered centroids, and at tiny sizes;
a scene finished with the device tree equals the generated scene."""
import numpy as np
import pytest

from specrel_amd import scene as S
from tests.test_extrude import SCENES, arrays, params, synthetic_boxes

pytestmark = pytest.mark.gpu


def assert_same_tree(got, ref, what):
    (gn, go), (rn, ro) = got[:2], ref[:2]
    assert gn.shape == rn.shape, f"{what}: {len(gn)} nodes vs {len(rn)}"
    same = gn.view(np.uint32) == rn.view(np.uint32)
    assert same.all(), f"{what}: {np.count_nonzero(~same.all(1))} nodes differ, first {np.nonzero(~same.all(1))[0][:5]}"
    assert np.array_equal(go, ro), f"{what}: leaf order differs at {np.count_nonzero(go != ro)} slots"


@pytest.mark.parametrize("name,over", SCENES)
def test_device_tree_equals_host_tree(gpu_ctx, name, over):
    d = S.DeferredScene(name, params(name, over))
    g, a, b, counts = S.extrude_tubes_host(d.tubes())
    got = gpu_ctx.build_bvh(b)
    assert_same_tree(got, S.build_bvh_host(b), name)
    d.set_records(g, a, b, counts)
    assert arrays(d.finish_with_tree(got[0], got[1])) == arrays(S.Scene.generate(name, params(name, over)))


def test_device_pipeline_extrusion_then_tree(gpu_ctx):
    p = S.gen_params("large", n_objects=24, slices=96, n_stars=10)
    d = S.DeferredScene("large", p)
    tubes = d.tubes()
    _, _, _, counts = gpu_ctx.extrude_tubes(tubes, download=False)
    n = int(counts.sum())
    got = gpu_ctx.build_bvh(None, n)  # the extrusion's records, still on the device
    g, a, b, c2 = S.extrude_tubes_host(tubes)
    assert_same_tree(got, S.build_bvh_host(b), "24 objects x 96 slices")
    st = got[2]
    print(f"device build of {n} records: {st}")
    assert st["levels"] > 3 and st["large_segments"] > 10 and st["small_subtrees"] > 100


@pytest.mark.parametrize("kind,n", [("uniform", 300000), ("clustered", 120000), ("clustered", 3000)])
def test_synthetic_bounds(gpu_ctx, kind, n):
    b = synthetic_boxes(11, n, kind)
    assert_same_tree(gpu_ctx.build_bvh(b), S.build_bvh_host(b), f"{kind} {n}")


@pytest.mark.parametrize("n", [1, 2, 3, 4, 31, 32, 33, 255, 256, 257, 5000, 40000])
def test_small_sizes(gpu_ctx, n):
    b = synthetic_boxes(5, n, "uniform")
    assert_same_tree(gpu_ctx.build_bvh(b), S.build_bvh_host(b), f"n={n}")
