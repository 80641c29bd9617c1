"""Wavefront form of the path tracer (persistent queue kernels, optional HIP
graph replay) against the megakernel and the oracle: identical bits for every
target, scene kind and launch shape."""
import numpy as np
import pytest

import pt_util as P
from mufflon_amd import hip
from mufflon_amd import scenes as S

pytestmark = pytest.mark.gpu


def _targets(ctx, pt, targets):
    return {t: ctx.pt_target(pt, t) for t in targets}


def _same(a, b):
    return all(np.array_equal(a[t].view(np.uint32), b[t].view(np.uint32)) for t in a)


@pytest.mark.parametrize("use_graph,tail", [(False, 0), (True, 0), (True, 1), (False, 3)])
def test_wavefront_equals_megakernel_material_zoo(hip_ctx, use_graph, tail):
    c = S.cornell_materials(12, 24, 40, 40)
    print("wavefront vs megakernel, material zoo, graph=%s, tail after %d" % (use_graph, tail), flush=True)
    targets = list(range(8))
    out = []
    for mode in (hip.PT_MEGAKERNEL, hip.PT_WAVEFRONT):
        pt, sc, bg = P.gpu_pt_materials(hip_ctx, c, "envmap")
        for t in targets:
            hip_ctx.pt_enable_target(pt, t, True)
        hip_ctx.pt_set_tuning(pt, tail_after=tail)
        hip_ctx.pt_set_mode(pt, mode, use_graph)
        hip_ctx.pt_iterate(pt, 3)
        assert hip_ctx.pt_iteration(pt) == 2
        out.append((_targets(hip_ctx, pt, targets), hip_ctx.pt_nan_count(pt)))
    for t in targets:
        print("target %d identical: %s" % (t, np.array_equal(out[0][0][t].view(np.uint32),
                                                              out[1][0][t].view(np.uint32))), flush=True)
    assert _same(out[0][0], out[1][0]) and out[0][1] == out[1][1]


def test_wavefront_matches_oracle_cornell(hip_ctx, oracle):
    c = P.cornell(16, 32, 48)
    print("wavefront vs oracle, cornell 48x48", flush=True)
    pt = P.gpu_pt(hip_ctx, c, nee_count=2, max_path_length=6)
    hip_ctx.pt_set_mode(pt, hip.PT_WAVEFRONT, True)
    hip_ctx.pt_iterate(pt, 5)
    ps = P.oracle_pt_scene(oracle, c, hip_ctx.pt_light_tree(pt))
    o_rad, o_hit, _ = oracle.pt_render(ps, 1, 0, 6, 2, 5)
    assert np.array_equal(hip_ctx.pt_target(pt, hip.TARGET_HIT_ID), o_hit)
    assert np.array_equal(hip_ctx.pt_target(pt).view(np.uint32), o_rad.view(np.uint32))


def test_wavefront_instanced_and_every_tuning(hip_ctx):
    """Instance level (TLAS traversal in the queue kernels) under every
    trace block / stack depth and two chunk sizes, plus reset and continue."""
    c = P.cornell_instanced(32)
    print("wavefront instanced, tuning sweep", flush=True)
    pt, sc = P.gpu_pt_instanced(hip_ctx, c)
    hip_ctx.pt_iterate(pt, 2)
    ref = hip_ctx.pt_target(pt)
    hip_ctx.pt_set_mode(pt, hip.PT_WAVEFRONT, True)
    shapes = [(64, 8), (64, 16), (128, 8), (128, 16), (256, 8), (256, 16), (256, 24)]
    for (block, depth), leaf in [(sh, 1) for sh in shapes] + [(sh, 8) for sh in shapes]:
        for chunk, sort, refill, tail in ((64, 0, 0, 0), (1024, 3, 0, 1), (256, 0, 16, 3), (256, 2, 3, 2)):
            hip_ctx.pt_set_tuning(pt, block, depth, chunk, sort, refill, leaf, tail)
            hip_ctx.pt_reset(pt)
            hip_ctx.pt_iterate(pt, 1)
            hip_ctx.pt_iterate(pt, 1)
            got = hip_ctx.pt_target(pt)
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (block, depth, leaf, chunk, sort, refill,
                                                                              tail)
    assert hip_ctx.pt_get_tuning(pt) == dict(traceBlock=256, stackDepth=24, queueChunk=256, sortBounces=2,
                                             refillBudget=3, leafBatch=8, tailAfter=2)
    # the megakernel with 8 triangles of a leaf loaded ahead
    hip_ctx.pt_set_mode(pt, hip.PT_MEGAKERNEL)
    hip_ctx.pt_reset(pt)
    hip_ctx.pt_iterate(pt, 2)
    assert np.array_equal(hip_ctx.pt_target(pt).view(np.uint32), ref.view(np.uint32))
    with pytest.raises(hip.HipError):
        hip_ctx.pt_set_tuning(pt, 256, 16, 256, 0, 0, 4)


def test_autotune_picks_a_shape_and_resets(hip_ctx):
    c = P.cornell(8, 16, 32)
    pt = P.gpu_pt(hip_ctx, c)
    hip_ctx.pt_set_mode(pt, hip.PT_WAVEFRONT, True)
    best, ms = hip_ctx.pt_autotune(pt, 1)
    print("autotune best %s %.3f ms" % (best, ms), flush=True)
    assert ms > 0 and best["queueChunk"] in (256, 1024) and best["sortBounces"] in (0, 1, 2)
    assert best["refillBudget"] in (0, 16, 32)
    assert hip_ctx.pt_iteration(pt) == -1
    hip_ctx.pt_iterate(pt, 2)
    a = hip_ctx.pt_target(pt)
    hip_ctx.pt_set_mode(pt, hip.PT_MEGAKERNEL)
    hip_ctx.pt_reset(pt)
    hip_ctx.pt_iterate(pt, 2)
    assert np.array_equal(a.view(np.uint32), hip_ctx.pt_target(pt).view(np.uint32))


def _with_stack(c, levels=160):
    """The Cornell box plus a stack of `levels` pairs of identical horizontal
    triangles 1e-4 extents apart: their boxes overlap so much that the SAH
    collapse keeps them in long leaves (the closest hit inside one is a
    tie between the two copies, won by the lower primitive index)."""
    box = P.scene_box(c)
    lo, hi = box[:3].astype(np.float64), box[3:].astype(np.float64)
    ext = float((hi - lo).max())
    ctr = 0.5 * (lo + hi)
    y = lo[1] + 0.3 * ext + np.arange(levels) * 1e-4 * ext
    tri = np.array([[-0.2, 0.0, -0.15], [0.2, 0.0, -0.15], [0.0, 0.0, 0.2]]) * ext
    v = np.repeat(tri[None], levels, 0) + np.stack([np.full(levels, ctr[0]), y, np.full(levels, ctr[2])], -1)[:, None]
    v = np.repeat(v, 2, 0).reshape(-1, 3).astype(np.float32)
    base = len(c["verts"])
    tris = (base + np.arange(len(v), dtype=np.uint32)).reshape(-1, 3)
    n = np.tile(np.array([0, 1, 0], np.float32), (len(v), 1))
    out = dict(c)
    out["verts"] = np.concatenate([c["verts"], v])
    out["normals"] = np.concatenate([c["normals"], n])
    out["uvs"] = np.concatenate([c["uvs"], np.zeros((len(v), 2), np.float32)])
    out["tris"] = np.concatenate([c["tris"], tris]).astype(c["tris"].dtype)
    out["mat_idx"] = np.concatenate([c["mat_idx"], np.full(len(tris), c["mat_idx"][0], c["mat_idx"].dtype)])
    return out


def test_leaf_batch_on_long_leaves(hip_ctx, oracle):
    """Leaves of up to hundreds of triangles (partial batches of 8, any-hit
    shadow rays that stop inside a batch, ties between identical triangles):
    megakernel and wavefront with leaf batch 1 and 8 give the oracle's bits."""
    c = _with_stack(P.cornell(8, 16, 48))
    print("long leaves: %d triangles" % len(c["tris"]), flush=True)
    lod = hip_ctx.lod_create(c["verts"], c["tris"])
    hip_ctx.lod_set_shading(lod, c["normals"], c["uvs"], c["mat_idx"])
    hip_ctx.lbvh_build(lod, P.scene_box(c))
    sc = hip_ctx.scene_create(np.array([lod], np.int32), S.IDENTITY34.reshape(1, 12))
    pt = hip_ctx.pt_create(sc, c["materials"], c["camera"],
                           dict(seed=1, min_path_length=0, max_path_length=6, nee_count=2))
    # the BVH really holds long leaves: 64-byte nodes, child (index, count) in dwords 6-7 and 14-15
    n_int, nodes, _ = hip_ctx.lbvh_get(lod, len(c["tris"]))
    q = nodes.view(np.int32).reshape(-1, 16)
    counts = np.concatenate([q[:, 7][q[:, 6] >= n_int], q[:, 15][q[:, 14] >= n_int]])
    print("largest leaf: %d triangles" % counts.max(), flush=True)
    assert counts.max() >= 64
    ps = P.oracle_pt_scene(oracle, c, hip_ctx.pt_light_tree(pt))
    o_rad, o_hit, _ = oracle.pt_render(ps, 1, 0, 6, 2, 3)
    for mode, leaf in ((hip.PT_MEGAKERNEL, 1), (hip.PT_MEGAKERNEL, 8), (hip.PT_WAVEFRONT, 8), (hip.PT_WAVEFRONT, 1)):
        hip_ctx.pt_set_mode(pt, mode, False)
        hip_ctx.pt_set_tuning(pt, 128, 8, 256, 0, 0, leaf)
        hip_ctx.pt_reset(pt)
        hip_ctx.pt_iterate(pt, 3)
        assert np.array_equal(hip_ctx.pt_target(pt, hip.TARGET_HIT_ID), o_hit), (mode, leaf)
        assert np.array_equal(hip_ctx.pt_target(pt).view(np.uint32), o_rad.view(np.uint32)), (mode, leaf)
