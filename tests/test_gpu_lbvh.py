"""GPU LBVH (through the C-ABI) against the oracle, bit for bit, at every
stage: Morton codes, sorted order, parents (topology), node boxes and counts,
the collapse scan and the final 64-byte-node BVH with its primitive ids."""
import json
import os

import numpy as np
import pytest

import lbvh_check as LC
import scene_util as U
from mufflon_amd import scenes as S

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _compare_build(hip_ctx, oracle, v, t, bbox=None):
    n = len(t)
    print("lbvh parity n=%d" % n, flush=True)
    ob = oracle.lbvh_build_tris(v, t, bbox)
    lod = hip_ctx.lod_create(v, t)
    g = hip_ctx.lbvh_build_stages(lod, n, ob["bbox"])
    assert np.array_equal(g["unsorted_keys"], ob["unsorted_keys"]), "morton"
    assert np.array_equal(g["sorted_keys"], ob["sorted_keys"]), "sorted keys"
    assert np.array_equal(g["sorted_ids"], ob["sorted_ids"]), "sorted ids"
    if n > 1:
        assert np.array_equal(g["parents"], ob["parents"]), "parents"
        gb, obx = g["boxes"].reshape(-1, 8), ob["boxes"].reshape(-1, 8)
        assert np.array_equal(gb[:, [0, 1, 2, 4, 5, 6]].view(np.uint32), obx[:, [0, 1, 2, 4, 5, 6]].view(np.uint32))
        assert np.array_equal(gb[:, 3].view(np.int32), obx[:, 3].view(np.int32)), "counts"
        # the SAH cost is consulted only below the 1023-primitive cap
        cnt = obx[:, 3].view(np.int32)
        small = np.array([not LC.lod_capped(int(c)) for c in cnt])
        assert np.array_equal(gb[small, 7].view(np.uint32), obx[small, 7].view(np.uint32)), "cost"
        assert np.array_equal(g["offsets"], ob["offsets"]), "collapse scan"
    nint, nodes, ids = hip_ctx.lbvh_get(lod, n)
    assert nint == ob["num_internal"]
    assert np.array_equal(nodes, ob["nodes"]), "collapsed nodes"
    assert np.array_equal(ids, ob["sorted_ids"])
    return lod, ob


@pytest.mark.parametrize("n,seed", [(1, 1), (2, 2), (3, 3), (257, 4), (2048, 14), (2049, 15), (4097, 5),
                                    (100_000, 6), (131_073, 16)])
def test_lbvh_matches_oracle(hip_ctx, oracle, n, seed):
    v, t = S.random_triangle_soup(n, seed=seed)
    _compare_build(hip_ctx, oracle, v, t)


@pytest.mark.parametrize("cap", [0, 16, 200])
def test_lbvh_level_capacity_fallback(hip_ctx, oracle, cap, monkeypatch):
    """Refit spans holding more straddling nodes than their LDS slots climb
    through global memory instead; every capacity builds the same BVH."""
    monkeypatch.setenv("MFH_LBVH_LEVEL_CAPACITY", str(cap))
    print("level capacity %d" % cap, flush=True)
    v, t = S.random_triangle_soup(300_000, seed=13)
    _compare_build(hip_ctx, oracle, v, t)


def test_lbvh_duplicates_and_identical(hip_ctx, oracle):
    v, t = S.random_triangle_soup(30_000, seed=8)
    t2 = np.concatenate([t, t[::5], t[::11]])
    _compare_build(hip_ctx, oracle, v, t2)
    v1, t1 = S.random_triangle_soup(1, seed=9)
    _compare_build(hip_ctx, oracle, v1, np.repeat(t1, 3000, axis=0))


def test_lbvh_golden_grid(hip_ctx, oracle):
    with open(os.path.join(GOLDEN, "lbvh_grid.json")) as f:
        gd = json.load(f)
    v = np.array(gd["vertices"], np.float32)
    t = np.array(gd["triangles"], np.uint32)
    lod = hip_ctx.lod_create(v, t)
    g = hip_ctx.lbvh_build_stages(lod, len(t), np.array(gd["bbox"], np.float32))
    assert [int(k) for k in g["unsorted_keys"]] == [int(k) for k in gd["morton"]]
    assert g["sorted_ids"].tolist() == gd["sorted_ids"]
    assert g["parents"].tolist() == gd["parents"]


def test_lbvh_repeatable(hip_ctx):
    v, t = S.random_triangle_soup(200_000, seed=12)
    lod = hip_ctx.lod_create(v, t)
    hip_ctx.lbvh_build(lod)
    a = hip_ctx.lbvh_get(lod, len(t))
    for _ in range(3):
        hip_ctx.lbvh_build(lod)
        b = hip_ctx.lbvh_get(lod, len(t))
        assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


@pytest.mark.slow
def test_lbvh_one_million_random_triangles(hip_ctx, oracle):
    v, t = S.random_triangle_soup(1_000_000, seed=7)
    _compare_build(hip_ctx, oracle, v, t)


def test_instance_level_bvh(hip_ctx, oracle):
    geoms, rng = U.small_lods(3)
    lods = []
    obj_boxes = []
    for v, t in geoms:
        lid = hip_ctx.lod_create(v, t)
        hip_ctx.lbvh_build(lid)
        lods.append(lid)
        obj_boxes.append(oracle.vertex_box(v))
    n_inst = 5000
    il = rng.integers(0, len(lods), n_inst).astype(np.int32)
    w = U.random_transforms(n_inst, 4, spread=50.0)
    sc = hip_ctx.scene_create(np.array(lods, np.int32)[il], w)
    box, nint, nodes, ids = hip_ctx.scene_get(sc, n_inst)
    ib, sb = oracle.instance_boxes(w, np.stack(obj_boxes), il)
    assert np.array_equal(box, sb)
    ob = oracle.lbvh_build_instances(w, np.stack(obj_boxes), il, sb)
    assert nint == ob["num_internal"]
    assert np.array_equal(ids, ob["sorted_ids"])
    assert np.array_equal(nodes, ob["nodes"])


@pytest.mark.slow
def test_lbvh_ten_million_terrain_triangles(hip_ctx, oracle):
    """The 2236 x 2236 fBm heightfield (9,990,450 triangles), every stage bit-exact."""
    hf = S.fbm_heightfield(2236, 8, 3)
    v, _, _, t = S.grid_mesh(hf, 1.0, 180.0)
    _compare_build(hip_ctx, oracle, v, t)
