"""The light tracer on the device (mfh_pt_lighttracer_*) against the oracle's
on the same seeds: Radiance, Lightness and Density means and variances.
The device adds its splats with float atomics in no fixed order, so the
bar is a tolerance: 1e-4 relative per value plus 1e-6 of the image's
largest value (sums of the same positive terms in another order).
Scenes: the Cornell box (one instance, smooth normals), the room of quads,
triangle boxes and spheres with quad, sphere and point lights, placed with
an instance level; a background that emits is refused."""
import numpy as np
import pytest

import pt_util as P
from mufflon_amd import hip
from mufflon_amd import scenes as S
from mufflon_amd import workloads as W

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def _close(g, o, what):
    tol = RTOL * np.abs(o) + 1e-6 * max(float(np.abs(o).max()), 1e-30)
    bad = np.abs(g - o) > tol
    print("%s: max |value| %.4g, max rel dev %.3g, out of tolerance %d" % (
        what, float(np.abs(o).max()), float((np.abs(g - o) / (np.abs(o) + 1e-30)).max()), int(bad.sum())), flush=True)
    assert not bad.any(), what


def _compare(hip_ctx, pt, ref, var):
    _close(hip_ctx.lighttracer_target(pt, hip.LT_RADIANCE), ref[..., :3], "radiance")
    _close(hip_ctx.lighttracer_target(pt, hip.LT_LIGHTNESS), ref[..., 3:4], "lightness")
    _close(hip_ctx.lighttracer_target(pt, hip.LT_DENSITY), ref[..., 4:5], "density")
    _close(hip_ctx.lighttracer_target(pt, hip.LT_RADIANCE, variance=True), var[..., :3], "radiance variance")
    assert ref[..., :3].mean() > 0


def test_lighttracer_cornell(hip_ctx, oracle):
    print("light tracer: cornell 40x40, depth 6", flush=True)
    c = P.cornell(12, 24, 40)
    pt = P.gpu_pt(hip_ctx, c, seed=3, max_path_length=6)
    hip_ctx.lighttracer_reset(pt, variance=True)
    hip_ctx.lighttracer_iterate(pt, 3, density=True)
    ps = P.oracle_pt_scene(oracle, c, light_tree=hip_ctx.pt_light_tree(pt))
    ref, var = oracle.lt_render(ps, 3, 0, 6, 3, density=True, variance=True)
    _compare(hip_ctx, pt, ref, var)
    # a reset starts over from the same seeds
    first = hip_ctx.lighttracer_target(pt, hip.LT_RADIANCE)
    hip_ctx.lighttracer_reset(pt, variance=True)
    hip_ctx.lighttracer_iterate(pt, 3, density=True)
    _close(hip_ctx.lighttracer_target(pt, hip.LT_RADIANCE), first, "radiance after a reset")


@pytest.mark.parametrize("instanced", [True, False], ids=["instanced", "single"])
def test_lighttracer_mixed_room(hip_ctx, oracle, instanced):
    print("light tracer: mixed room, instanced %s" % instanced, flush=True)
    sc = S.mixed_scene(48, 40, instanced=instanced)
    pt, scid, _ = W.path_tracer(hip_ctx, sc, seed=2, max_path_length=5)
    hip_ctx.lighttracer_reset(pt, variance=True)
    hip_ctx.lighttracer_iterate(pt, 2, density=True)
    box, _, _, _ = hip_ctx.scene_get(scid, len(sc["inst_lod"]))
    ps = P.oracle_pt_scene_general(oracle, sc, hip_ctx.pt_light_tree(pt), box)
    ref, var = oracle.lt_render(ps, 2, 0, 5, 2, density=True, variance=True)
    _compare(hip_ctx, pt, ref, var)


def test_lighttracer_refuses_an_emitting_background(hip_ctx):
    c = S.cornell_materials(8, 16, 16, 16)
    pt, _, _ = P.gpu_pt_materials(hip_ctx, c, background="envmap")
    with pytest.raises(hip.HipError):
        hip_ctx.lighttracer_iterate(pt, 1)


def test_lighttracer_refuses_lights_that_emit_nothing(hip_ctx):
    c = P.cornell(4, 8, 8)
    mats = c["materials"].copy()
    mats["radiance"][mats["type"] == S.MAT_EMISSIVE] = 0.0
    pt = P.gpu_pt(hip_ctx, dict(c, materials=mats), max_path_length=4)
    with pytest.raises(hip.HipError):
        hip_ctx.lighttracer_iterate(pt, 1)
