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
    for chunk, refill, leaf in ((1024, 0, 1), (256, 16, 1), (256, 5, 1), (256, 0, 8), (1024, 16, 8)):
        hip_ctx.batch_set_tuning(b, 128, 16, chunk, refill, leaf)
        hip_ctx.batch_reset(b)
        hip_ctx.batch_iterate(b, 3)
        for pt, r in zip(pts, ref):
            assert np.array_equal(hip_ctx.pt_target(pt, 0).view(np.uint32), r[0].view(np.uint32)), (chunk, refill, leaf)
    with pytest.raises(hip.HipError):
        hip_ctx.pt_iterate(pts[0], 1)
        hip_ctx.batch_iterate(b, 1)
