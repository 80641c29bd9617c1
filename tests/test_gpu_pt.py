"""GPU path tracer (through the C-ABI) against the oracle path tracer with
identical seeds: per-pixel running means of Radiance and Hit ID after N
iterations. Tolerance: 1e-4 relative per pixel (plus 1e-6 absolute for
near-black pixels); the kernels evaluate the same IEEE operations in the same
order, so the images are expected to agree bit for bit, which is reported."""
import numpy as np
import pytest

import pt_util as P
from mufflon_amd import hip
from mufflon_amd import scenes as S

pytestmark = pytest.mark.gpu

RTOL = 1e-4
ATOL = 1e-6


def _compare(hip_ctx, oracle, c, iterations, seed=1, max_path_length=8, nee_count=1):
    print("pt parity %dx%d it=%d depth=%d nee=%d" % (c["camera"]["width"], c["camera"]["height"], iterations,
                                                    max_path_length, nee_count), flush=True)
    pt = P.gpu_pt(hip_ctx, c, seed=seed, max_path_length=max_path_length, nee_count=nee_count)
    hip_ctx.pt_iterate(pt, iterations)
    g_rad = hip_ctx.pt_target(pt, hip.TARGET_RADIANCE)
    g_hit = hip_ctx.pt_target(pt, hip.TARGET_HIT_ID)
    assert hip_ctx.pt_iteration(pt) == iterations - 1
    # the oracle consumes the very light tree the device path built
    lt = hip_ctx.pt_light_tree(pt)
    ps = P.oracle_pt_scene(oracle, c, lt)
    o_rad, o_hit, _ = oracle.pt_render(ps, seed, 0, max_path_length, nee_count, iterations)
    assert np.array_equal(g_hit, o_hit), "hit ids"
    diff = np.abs(g_rad - o_rad)
    bad = diff > RTOL * np.abs(o_rad) + ATOL
    exact = np.array_equal(g_rad.view(np.uint32), o_rad.view(np.uint32))
    print("radiance: bit-exact=%s, max rel diff=%.3g, pixels out of tolerance=%d" %
          (exact, float((diff / np.maximum(np.abs(o_rad), 1e-30)).max()), int(bad.any(axis=2).sum())), flush=True)
    assert not bad.any(), "per-pixel radiance differs beyond 1e-4 relative"
    return g_rad, o_rad


def test_pt_small_cornell(hip_ctx, oracle):
    c = P.cornell(16, 32, 64)
    rad, _ = _compare(hip_ctx, oracle, c, 8)
    assert rad.mean() > 0.01


def test_pt_nee_count_and_depth(hip_ctx, oracle):
    c = P.cornell(12, 24, 48)
    _compare(hip_ctx, oracle, c, 4, seed=3, max_path_length=3, nee_count=2)


def test_pt_reset_restarts_sequence(hip_ctx):
    c = P.cornell(8, 16, 32)
    pt = P.gpu_pt(hip_ctx, c)
    hip_ctx.pt_iterate(pt, 3)
    a = hip_ctx.pt_target(pt)
    hip_ctx.pt_reset(pt)
    assert hip_ctx.pt_iteration(pt) == -1
    hip_ctx.pt_iterate(pt, 3)
    b = hip_ctx.pt_target(pt)
    assert np.array_equal(a, b)


@pytest.mark.slow
def test_pt_cornell_256_64_iterations(hip_ctx, oracle):
    """The Cornell box workload: 256x256, 128x256-segment sphere, depth 8, 64 iterations, seed 1."""
    c = P.cornell(128, 256, 256)
    _compare(hip_ctx, oracle, c, 64)


def test_pt_instanced_scene(hip_ctx, oracle):
    """Instance level in the path tracer: transformed LoDs, hit frames through
    the instance transforms, area lights gathered in world space."""
    c = P.cornell_instanced(48)
    print("pt parity instanced scene, %d instances" % len(c["inst_lod"]), flush=True)
    pt, sc = P.gpu_pt_instanced(hip_ctx, c)
    hip_ctx.pt_iterate(pt, 6)
    g_rad = hip_ctx.pt_target(pt, hip.TARGET_RADIANCE)
    g_hit = hip_ctx.pt_target(pt, hip.TARGET_HIT_ID)
    box, _, _, _ = hip_ctx.scene_get(sc, len(c["inst_lod"]))
    lt = hip_ctx.pt_light_tree(pt)
    assert lt[0]["posLightCount"] == 2
    ps = P.oracle_pt_scene_instanced(oracle, c, lt, box)
    o_rad, o_hit, _ = oracle.pt_render(ps, 1, 0, 8, 1, 6)
    assert np.array_equal(g_hit, o_hit)
    diff = np.abs(g_rad - o_rad)
    print("instanced radiance: bit-exact=%s max abs diff=%.3g" %
          (np.array_equal(g_rad.view(np.uint32), o_rad.view(np.uint32)), float(diff.max())), flush=True)
    assert not (diff > RTOL * np.abs(o_rad) + ATOL).any()
    assert g_rad.mean() > 0.01


@pytest.mark.parametrize("background", ["none", "colored", "envmap"])
def test_pt_material_zoo(hip_ctx, oracle, background):
    """TORRANCE (GGX V-cavity with a roughness texture, anisotropic GGX Smith,
    Beckmann), absorbing MICROFACET glass, a linear albedo texture, point
    lights, the ceiling light and each background, against the oracle."""
    c = S.cornell_materials(16, 32, 48, 48)
    print("pt parity material zoo, background %s" % background, flush=True)
    pt, sc, bg = P.gpu_pt_materials(hip_ctx, c, background)
    hip_ctx.pt_iterate(pt, 6)
    g_rad = hip_ctx.pt_target(pt, hip.TARGET_RADIANCE)
    g_hit = hip_ctx.pt_target(pt, hip.TARGET_HIT_ID)
    box, _, _, _ = hip_ctx.scene_get(sc, 1)
    lt = hip_ctx.pt_light_tree(pt)
    ps = P.oracle_pt_scene_materials(oracle, c, bg, light_tree=lt, box=box)
    o_rad, o_hit, o_nans = oracle.pt_render(ps, 1, 0, 8, 1, 6)
    assert np.array_equal(g_hit, o_hit)
    diff = np.abs(g_rad - o_rad)
    bad = diff > RTOL * np.abs(o_rad) + ATOL
    print("material zoo %s: bit-exact=%s, pixels out of tolerance=%d, max abs diff=%.3g" %
          (background, np.array_equal(g_rad.view(np.uint32), o_rad.view(np.uint32)), int(bad.any(axis=2).sum()),
           float(diff.max())), flush=True)
    assert not bad.any()
    assert hip_ctx.pt_nan_count(pt) == o_nans


def test_pt_all_targets(hip_ctx, oracle):
    """Every PtTargets target (Radiance, Hit ID, Position, Depth, Normal, UV,
    Albedo, Lightness) against the oracle, with textures and an env map."""
    c = S.cornell_materials(12, 24, 40, 40)
    print("pt parity all targets", flush=True)
    pt, sc, bg = P.gpu_pt_materials(hip_ctx, c, "envmap")
    for t in range(8):
        hip_ctx.pt_enable_target(pt, t, True)
    hip_ctx.pt_iterate(pt, 4)
    box, _, _, _ = hip_ctx.scene_get(sc, 1)
    ps = P.oracle_pt_scene_materials(oracle, c, bg, light_tree=hip_ctx.pt_light_tree(pt), box=box)
    ref, nans = oracle.pt_render_targets(ps, 1, 0, 8, 1, 4, list(range(8)))
    for t in range(8):
        g = hip_ctx.pt_target(pt, t)
        o = ref[t]
        bad = np.abs(g - o) > RTOL * np.abs(o) + ATOL
        print("target %d: bit-exact=%s, out of tolerance=%d" %
              (t, np.array_equal(g.view(np.uint32), o.view(np.uint32)), int(bad.sum())), flush=True)
        assert not bad.any(), t
    assert hip_ctx.pt_nan_count(pt) == nans
    hip_ctx.pt_enable_target(pt, hip.TARGET_DEPTH, False)
    with pytest.raises(hip.HipError):
        hip_ctx.pt_target(pt, hip.TARGET_DEPTH)


@pytest.mark.parametrize("mode", [hip.PT_MEGAKERNEL, hip.PT_WAVEFRONT])
def test_pt_target_variance(hip_ctx, oracle, mode):
    """Running variance of Radiance and Depth (render_enable_render_target with
    variance) against the oracle, megakernel and wavefront forms; disabling
    the variance keeps the mean."""
    c = S.cornell_materials(12, 24, 40, 40)
    print("pt parity variance mode=%d" % mode, flush=True)
    pt, sc, bg = P.gpu_pt_materials(hip_ctx, c, "envmap")
    hip_ctx.pt_set_mode(pt, mode, mode == hip.PT_WAVEFRONT)
    hip_ctx.pt_enable_variance(pt, hip.TARGET_RADIANCE, True)
    hip_ctx.pt_enable_variance(pt, hip.TARGET_DEPTH, True)
    hip_ctx.pt_iterate(pt, 6)
    box, _, _, _ = hip_ctx.scene_get(sc, 1)
    ps = P.oracle_pt_scene_materials(oracle, c, bg, light_tree=hip_ctx.pt_light_tree(pt), box=box)
    mean, var, _ = oracle.pt_render_targets(ps, 1, 0, 8, 1, 6, [0, 1, 3], variance=[0, 3])
    for t in (0, 3):
        for g, o, what in ((hip_ctx.pt_target(pt, t), mean[t], "mean"), (hip_ctx.pt_variance(pt, t), var[t], "var")):
            bad = np.abs(g - o) > RTOL * np.abs(o) + ATOL
            print("target %d %s: bit-exact=%s, out of tolerance=%d" %
                  (t, what, np.array_equal(g.view(np.uint32), o.view(np.uint32)), int(bad.sum())), flush=True)
            assert not bad.any(), (t, what)
    with pytest.raises(hip.HipError):
        hip_ctx.pt_variance(pt, hip.TARGET_HIT_ID)
    hip_ctx.pt_enable_variance(pt, hip.TARGET_DEPTH, False)
    hip_ctx.pt_iterate(pt, 1)
    hip_ctx.pt_target(pt, hip.TARGET_DEPTH)
    with pytest.raises(hip.HipError):
        hip_ctx.pt_variance(pt, hip.TARGET_DEPTH)


@pytest.mark.parametrize("mode", [hip.PT_MEGAKERNEL, hip.PT_WAVEFRONT])
def test_pt_nee_position_guide(hip_ctx, oracle, mode):
    """"NEE position guide" (distance-weighted light tree traversal in NEE and
    in the hit MIS) against the oracle: area and point lights plus an env
    map, megakernel and wavefront; changing the parameter resets and changes
    the image."""
    c = S.cornell_materials(12, 24, 40, 40)
    print("pt parity position guide mode=%d" % mode, flush=True)
    pt, sc, bg = P.gpu_pt_materials(hip_ctx, c, "envmap", pos_guide=True, nee_count=2)
    hip_ctx.pt_set_mode(pt, mode, mode == hip.PT_WAVEFRONT)
    hip_ctx.pt_iterate(pt, 4)
    g = hip_ctx.pt_target(pt)
    box, _, _, _ = hip_ctx.scene_get(sc, 1)
    lt = hip_ctx.pt_light_tree(pt)
    ps = P.oracle_pt_scene_materials(oracle, c, bg, light_tree=lt, box=box, pos_guide=True)
    o, _ = oracle.pt_render_targets(ps, 1, 0, 8, 2, 4, [0, 1])
    o = o[0]
    bad = np.abs(g - o) > RTOL * np.abs(o) + ATOL
    print("position guide: bit-exact=%s, out of tolerance=%d" %
          (np.array_equal(g.view(np.uint32), o.view(np.uint32)), int(bad.any(axis=2).sum())), flush=True)
    assert not bad.any()
    hip_ctx.pt_set_params(pt, dict(seed=1, max_path_length=8, nee_count=2, nee_position_guide=False))
    assert hip_ctx.pt_iteration(pt) == -1
    hip_ctx.pt_iterate(pt, 4)
    g2 = hip_ctx.pt_target(pt)
    ps2 = P.oracle_pt_scene_materials(oracle, c, bg, light_tree=lt, box=box, pos_guide=False)
    o2, _ = oracle.pt_render_targets(ps2, 1, 0, 8, 2, 4, [0, 1])
    assert not (np.abs(g2 - o2[0]) > RTOL * np.abs(o2[0]) + ATOL).any()
    assert not np.array_equal(g, g2)
