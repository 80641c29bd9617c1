"""mfh_lbvh_build_noalloc: the LBVH build that allocates nothing (scratch in a
static device array per GPU, caller-owned inputs and outputs) against the
existing build (mfh_lbvh_build + mfh_lbvh_get) on the same LoD, bit for bit,
on consecutive calls of decreasing size; size and index errors."""
import numpy as np
import pytest

from mufflon_amd import scenes as S

pytestmark = pytest.mark.gpu

MAX_TRIS = 262144  # MFH_NOALLOC_MAX_TRIANGLES


def _reference_build(ctx, v, t, box):
    lod = ctx.lod_create(v, t)
    ctx.lbvh_build(lod, box)
    return ctx.lbvh_get(lod, len(t))


def _noalloc_build(ctx, v, t, box):
    dv = ctx.to_device(np.ascontiguousarray(v, np.float32))
    dt = ctx.to_device(np.ascontiguousarray(t, np.uint32))
    dn = ctx.alloc(64 * max(len(t) - 1, 1))
    di = ctx.alloc(4 * len(t))
    try:
        nint = ctx.lbvh_build_noalloc(dv.ptr, len(v), dt.ptr, len(t), box, dn.ptr, di.ptr)
        nodes = dn.download(np.uint8, 64 * nint)
        ids = di.download(np.int32, len(t))
    finally:
        for b in (dv, dt, dn, di):
            b.free()
    return nint, nodes, ids


def _box(v):
    return np.concatenate([v.min(0), v.max(0)]).astype(np.float32)


def _check_same(ctx, v, t):
    box = _box(v)
    print("noalloc build: %d triangles" % len(t), flush=True)
    rn, rnodes, rids = _reference_build(ctx, v, t, box)
    nn, nnodes, nids = _noalloc_build(ctx, v, t, box)
    assert nn == rn
    assert np.array_equal(nnodes, rnodes)
    assert np.array_equal(nids, rids)
    return rn


def test_noalloc_build_matches_build_on_consecutive_calls():
    from mufflon_amd.hip import HipContext
    with HipContext(0) as ctx:
        # the largest LoD it takes, then smaller ones: the array keeps the
        # previous build's contents, which must not leak into the next
        v, t = S.random_triangle_soup(MAX_TRIS, seed=21)
        assert _check_same(ctx, v, t) > 1000
        v2, t2 = S.random_triangle_soup(5000, seed=22, mu=-4.0)
        assert _check_same(ctx, v2, t2) > 10
        # duplicate centroids (equal Morton codes) and a tiny LoD
        v3 = np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (64, 1))
        t3 = np.arange(3 * 64, dtype=np.uint32).reshape(-1, 3)
        _check_same(ctx, v3, t3)
        # one triangle: no internal nodes
        nn, _, ids = _noalloc_build(ctx, v3[:3], t3[:1], _box(v3[:3]))
        assert nn == 0 and ids.tolist() == [0]


def test_noalloc_build_rejects_what_it_cannot_hold():
    from mufflon_amd.hip import HipContext, HipError
    with HipContext(0) as ctx:
        v, t = S.random_triangle_soup(1000, seed=23)
        dv = ctx.to_device(v)
        dt = ctx.to_device(t)
        dn = ctx.alloc(64 * 1000)
        di = ctx.alloc(4 * 1000)
        # more triangles than the scratch holds: refused on the host, before any launch
        with pytest.raises(HipError) as e:
            ctx.lbvh_build_noalloc(dv.ptr, len(v), dt.ptr, MAX_TRIS + 1, _box(v), dn.ptr, di.ptr)
        assert e.value.code == 1 and "at most" in str(e.value)
        # a vertex index past the vertex count (only the indices are read)
        with pytest.raises(HipError) as e:
            ctx.lbvh_build_noalloc(dv.ptr, len(v) - 1, dt.ptr, len(t), _box(v), dn.ptr, di.ptr)
        assert e.value.code == 1 and "index" in str(e.value)
        # the array is usable after the refusals
        nint = ctx.lbvh_build_noalloc(dv.ptr, len(v), dt.ptr, len(t), _box(v), dn.ptr, di.ptr)
        lod = ctx.lod_create(v, t)
        ctx.lbvh_build(lod, _box(v))
        rn, rnodes, rids = ctx.lbvh_get(lod, len(t))
        assert nint == rn
        assert np.array_equal(dn.download(np.uint8, 64 * nint), rnodes)
        assert np.array_equal(di.download(np.int32, len(t)), rids)
