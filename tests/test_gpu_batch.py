"""Batches of scenes in one wavefront launch per stage: every scene's
targets equal the same path tracer run alone (megakernel), whatever the mix of
resolutions, path lengths, instance levels, materials and backgrounds."""
import numpy as np
import pytest

import pt_util as P
from mufflon_amd import hip
from mufflon_amd import scenes as S

pytestmark = pytest.mark.gpu


def _make(hip_ctx):
    pts = []
    c1 = P.cornell(12, 24, 40)
    pts.append(P.gpu_pt(hip_ctx, c1, max_path_length=5, nee_count=2))
    c2 = S.cornell_materials(8, 16, 24, 24)
    pt2, _, _ = P.gpu_pt_materials(hip_ctx, c2, "envmap", max_path_length=8)
    hip_ctx.pt_enable_target(pt2, hip.TARGET_NORMAL, True)
    pts.append(pt2)
    c3 = P.cornell_instanced(32)
    pt3, _ = P.gpu_pt_instanced(hip_ctx, c3, max_path_length=12)
    pts.append(pt3)
    return pts


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_equals_single_scene_runs(hip_ctx, use_graph):
    print("batch of 3 scenes, graph=%s" % use_graph, flush=True)
    pts = _make(hip_ctx)
    ref = []
    for pt in pts:
        hip_ctx.pt_iterate(pt, 3)
        ref.append([hip_ctx.pt_target(pt, t) for t in (0, 1)])
    b = hip_ctx.batch_create(pts, use_graph)
    hip_ctx.batch_iterate(b, 2)
    hip_ctx.batch_iterate(b, 1)
    for pt, r in zip(pts, ref):
        assert hip_ctx.pt_iteration(pt) == 2
        for t, want in zip((0, 1), r):
            got = hip_ctx.pt_target(pt, t)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (pt, t)
    for chunk, refill, leaf, tail in ((1024, 0, 1, 0), (256, 16, 1, 0), (256, 5, 1, 2), (256, 0, 8, 1),
                                      (1024, 16, 8, 0), (256, 0, 8, 3)):
        hip_ctx.batch_set_tuning(b, 128, 16, chunk, refill, leaf, tail)
        hip_ctx.batch_reset(b)
        hip_ctx.batch_iterate(b, 3)
        for pt, r in zip(pts, ref):
            assert np.array_equal(hip_ctx.pt_target(pt, 0).view(np.uint32), r[0].view(np.uint32)), (chunk, refill, leaf,
                                                                                                    tail)
    with pytest.raises(hip.HipError):
        hip_ctx.pt_iterate(pts[0], 1)
        hip_ctx.batch_iterate(b, 1)


def test_uneven_batch_48_scenes(hip_ctx, oracle):
    """The uneven-batch benchmark workload: 48 procedural scenes (seeds
    100-147; 10k-2M triangles, 128^2-1024^2 pixels, maximum path length
    4-16, env maps, point lights) in one HIP-graph-replayed batch, autotuned.
    Every scene's Radiance equals the same path tracer run alone (megakernel)
    bit for bit; scenes 0-2 also equal the oracle on a 32x32 window; the
    albedo-texture gradients of scenes 0-7 (adjoint image nonzero on a 32x32
    window) equal the oracle's FP64 adjoint."""
    from mufflon_amd import workloads as W
    its = 2
    scenes = [S.batch_scene(i) for i in range(48)]
    print("uneven batch: %d scenes, %d triangles" % (len(scenes), sum(len(s["lods"][0]["tris"]) for s in scenes)),
          flush=True)
    pts, texs, scids = [], [], []
    for sc in scenes:
        pt, scid, tex = W.path_tracer(hip_ctx, sc, max_path_length=sc["max_path_length"])
        pts.append(pt)
        texs.append(tex)
        scids.append(scid)
    ref = []
    for pt in pts:
        hip_ctx.pt_iterate(pt, its)
        ref.append(hip_ctx.pt_target(pt))
    b = hip_ctx.batch_create(pts, True)
    tuned, _ = hip_ctx.batch_autotune(b, 1)
    print("batch tuning", tuned, flush=True)
    hip_ctx.batch_reset(b)
    hip_ctx.batch_iterate(b, its)
    for k, (pt, want) in enumerate(zip(pts, ref)):
        got = hip_ctx.pt_target(pt)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), "scene %d" % k
    for k in range(3):
        sc = scenes[k]
        w, h = sc["camera"]["width"], sc["camera"]["height"]
        x0, y0 = w // 2 - 16, h // 2 - 16
        box, _, _, _ = hip_ctx.scene_get(scids[k], 1)
        ps = P.oracle_pt_scene_general(oracle, sc, hip_ctx.pt_light_tree(pts[k]), box)
        win = oracle.pt_render_window(ps, 1, 0, sc["max_path_length"], 1, its, x0, y0, 32, 32)
        same = np.array_equal(ref[k][y0:y0 + 32, x0:x0 + 32].view(np.uint32), win.view(np.uint32))
        print("scene %d (%dx%d, %d triangles): oracle window bit-exact %s" %
              (k, w, h, len(sc["lods"][0]["tris"]), same), flush=True)
        assert same
    for k in range(8):
        sc = scenes[k]
        w, h = sc["camera"]["width"], sc["camera"]["height"]
        delta = np.zeros((h, w, 3), np.float32)
        x0, y0 = w // 2 - 16, h // 2 - 16
        delta[y0:y0 + 32, x0:x0 + 32] = np.random.default_rng(k).normal(size=(32, 32, 3))
        hip_ctx.pt_adjoint(pts[k], its, delta)
        g = hip_ctx.pt_texture_gradient(pts[k], int(texs[k][1]), (64, 64, 3))
        box, _, _, _ = hip_ctx.scene_get(scids[k], 1)
        ps = P.oracle_pt_scene_general(oracle, sc, hip_ctx.pt_light_tree(pts[k]), box)
        go, _ = oracle.pt_adjoint(ps, 1, 0, sc["max_path_length"], 1, its, delta, {1: (64, 64)})
        go = go[1][..., :3]
        scale = np.abs(go).max()
        err = np.abs(g - go)
        print("scene %d albedo gradient: max |grad| %.4g, max abs err %.3g" % (k, scale, float(err.max())),
              flush=True)
        assert (err <= 1e-3 * np.abs(go) + 1e-5 * scale).all()
