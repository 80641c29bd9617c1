"""Pins the LBVH oracle (oracle/lbvh_oracle.cpp) with checks that do not
restate its code: naive bit interleaving, exact dyadic Morton inputs, a stable
numpy argsort, brute-force Karras topology, exact AABB unions, the collapse
characterization, structural invariants of the collapsed BVH and a committed
known-answer fixture generated without the oracle."""
import json
import os

import numpy as np
import pytest

import lbvh_check as LC
from mufflon_amd import scenes as S

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_part_by_two21_all_inputs(oracle):
    x = np.arange(1 << 21, dtype=np.uint32)
    assert np.array_equal(oracle.part_by_two21(x), LC.spread21(x))


def test_morton_exact_dyadic_points(oracle):
    rng = np.random.default_rng(1)
    q = rng.integers(0, 1 << 21, size=(5000, 3)).astype(np.uint64)
    pts = (q.astype(np.float64) / float(1 << 21)).astype(np.float32)  # exact in float32
    bbox = np.array([0, 0, 0, 1, 1, 1], np.float32)
    codes = oracle.morton_points(pts, bbox)
    assert np.array_equal(codes, LC.interleave21((q[:, 0], q[:, 1], q[:, 2])))


def test_morton_clamps_and_corners(oracle):
    bbox = np.array([-2, -2, -2, 2, 2, 2], np.float32)
    pts = np.array([[-2, -2, -2], [2, 2, 2], [-5, 0, 9], [0, 0, 0]], np.float32)
    codes = oracle.morton_points(pts, bbox)
    assert codes[0] == 0
    assert codes[1] == (1 << 63) - 1  # every one of the 63 bits set
    top = (1 << 21) - 1
    half = 1 << 20
    assert codes[2] == LC.interleave21(([0], [half], [top]))[0]
    assert codes[3] == LC.interleave21(([half], [half], [half]))[0]
    # zero span on an axis: sdiv divides by 1e-20, so the point on the plane
    # maps to 0 and a point one ulp above it clamps to the top
    flat = np.array([0, 0, 1, 1, 1, 1], np.float32)
    pts2 = np.array([[0.5, 0.5, 1.0], [0.5, 0.5, np.nextafter(np.float32(1), np.float32(2))]], np.float32)
    c = oracle.morton_points(pts2, flat)
    assert c[0] == LC.interleave21(([half], [half], [0]))[0]
    assert c[1] == LC.interleave21(([half], [half], [top]))[0]


def _soup_with_duplicates(n, seed):
    v, t = S.random_triangle_soup(n, seed=seed)
    # duplicate a few triangles so equal Morton codes occur
    dup = np.arange(0, n, 7)
    t2 = np.concatenate([t, t[dup]])
    return v, t2


def test_sort_is_stable_by_id(oracle):
    v, t = _soup_with_duplicates(4000, 3)
    b = oracle.lbvh_build_tris(v, t)
    order = np.argsort(b["unsorted_keys"], kind="stable")
    assert np.array_equal(b["sorted_ids"], order.astype(np.int32))
    assert np.array_equal(b["sorted_keys"], b["unsorted_keys"][order])


@pytest.mark.parametrize("n", [2, 3, 5, 17, 64, 257, 1000])
def test_karras_matches_bruteforce(oracle, n):
    rng = np.random.default_rng(n)
    keys = np.sort(rng.integers(0, 1 << 12, n).astype(np.uint64))  # many duplicates
    assert np.array_equal(oracle.karras(keys), LC.karras_bruteforce(keys))


def test_karras_all_equal_keys(oracle):
    keys = np.full(33, 12345, np.uint64)
    assert np.array_equal(oracle.karras(keys), LC.karras_bruteforce(keys))


def _check_refit(b, verts, tris):
    n = len(tris)
    parents = b["parents"]
    B = b["boxes"].reshape(-1, 8)
    left, right = LC.children_from_parents(parents, n)
    tlo, thi = LC.tri_boxes(verts, tris)
    ids = b["sorted_ids"]
    nint = n - 1
    # leaves: exact triangle boxes, count 1, cost 2 + 1.2
    assert np.array_equal(B[nint:, 0:3], tlo[ids])
    assert np.array_equal(B[nint:, 4:7], thi[ids])
    assert (B[nint:, 3].view(np.int32) == (1 << 20)).all()
    assert (B[nint:, 7] == np.float32(2.0) + np.float32(1.2)).all()
    # internal: exact union of the children, summed (saturating) counts
    L, R = B[left], B[right]
    assert np.array_equal(B[:nint, 0:3], np.minimum(L[:, 0:3], R[:, 0:3]))
    assert np.array_equal(B[:nint, 4:7], np.maximum(L[:, 4:7], R[:, 4:7]))
    cl, cr = L[:, 3].view(np.int32) >> 20, R[:, 3].view(np.int32) >> 20
    assert np.array_equal(B[:nint, 3].view(np.int32) >> 20, np.minimum(cl + cr, 1023))
    # leaf count below a node is the size of its key range
    return B, L, R


def _surface(lo, hi):
    d = (hi - lo).astype(np.float32)
    return np.float32(2.0) * ((d[:, 0] * d[:, 1] + d[:, 1] * d[:, 2]) + d[:, 2] * d[:, 0])


def test_refit_union_counts_and_cost_order(oracle):
    v, t = S.random_triangle_soup(3000, seed=11)
    b = oracle.lbvh_build_tris(v, t)
    B, L, R = _check_refit(b, v, t)
    nint = len(t) - 1
    cnt = B[:nint, 3].view(np.int32)
    small = (cnt >> 20) < 1023
    sa = _surface(B[:nint, 0:3], B[:nint, 4:7])
    saL = _surface(L[:, 0:3], L[:, 4:7])
    saR = _surface(R[:, 0:3], R[:, 4:7])
    cL = np.float32(2.0) + (L[:, 3].view(np.int32) >> 20).astype(np.float32) * np.float32(1.2)
    cR = np.float32(2.0) + (R[:, 3].view(np.int32) >> 20).astype(np.float32) * np.float32(1.2)
    # right operand first: the serial reference completes a node from its right subtree
    cost = (np.float32(2.0) + (saR / sa) * cR) + (saL / sa) * cL
    assert np.array_equal(B[:nint, 7][small], cost[small])


def test_collapse_marks_match_per_node_rule(oracle):
    v, t = S.random_triangle_soup(5000, seed=5)
    b = oracle.lbvh_build_tris(v, t)
    n = len(t)
    offs = b["offsets"]
    marks = np.diff(np.concatenate([[0], offs]))
    exp = LC.expected_marks(b["parents"], b["boxes"], n, LC.lod_capped, LC.lod_leaf_cost)
    assert np.array_equal(marks, exp)
    assert b["num_internal"] == offs[-1]


@pytest.mark.parametrize("n,seed", [(2, 1), (3, 2), (10, 3), (1000, 4), (20000, 6)])
def test_collapsed_bvh_invariants(oracle, n, seed):
    v, t = S.random_triangle_soup(n, seed=seed)
    b = oracle.lbvh_build_tris(v, t)
    nint = b["num_internal"]
    assert nint >= 1
    leaves = LC.walk_collapsed(b["nodes"], nint, n)
    starts = sorted((s, c) for s, c, _, _ in leaves)
    pos = 0
    for s, c in starts:
        assert s == pos and 1 <= c <= 1022
        pos += c
    assert pos == n
    tlo, thi = LC.tri_boxes(v, t)
    ids = b["sorted_ids"]
    for s, c, lo, hi in leaves:
        p = ids[s:s + c]
        assert (lo <= tlo[p].min(axis=0)).all() and (hi >= thi[p].max(axis=0)).all()


def test_single_triangle_has_no_hierarchy(oracle):
    v, t = S.random_triangle_soup(1, seed=1)
    b = oracle.lbvh_build_tris(v, t)
    assert b["num_internal"] == 0 and b["sorted_ids"][0] == 0


def test_identical_triangles(oracle):
    v, t = S.random_triangle_soup(1, seed=2)
    t = np.repeat(t, 40, axis=0)
    b = oracle.lbvh_build_tris(v, t)
    assert (b["sorted_keys"] == b["sorted_keys"][0]).all()
    assert np.array_equal(b["sorted_ids"], np.arange(40))
    assert np.array_equal(b["parents"], LC.karras_bruteforce(b["sorted_keys"]))
    # all 40 identical prims collapse into one leaf under the root
    leaves = LC.walk_collapsed(b["nodes"], b["num_internal"], 40)
    assert sum(c for _, c, _, _ in leaves) == 40


def test_golden_grid_scene(oracle):
    with open(os.path.join(GOLDEN, "lbvh_grid.json")) as f:
        g = json.load(f)
    v = np.array(g["vertices"], np.float32)
    t = np.array(g["triangles"], np.uint32)
    b = oracle.lbvh_build_tris(v, t, np.array(g["bbox"], np.float32))
    assert [int(k) for k in b["unsorted_keys"]] == [int(k) for k in g["morton"]]
    assert b["sorted_ids"].tolist() == g["sorted_ids"]
    assert b["parents"].tolist() == g["parents"]


def test_instance_level_build(oracle):
    rng = np.random.default_rng(9)
    n = 300
    lod_boxes = np.array([[0, 0, 0, 1, 1, 1], [-1, -2, 0, 1, 2, 0.5], [0, 0, 0, 0.1, 0.1, 0.1]], np.float32)
    il = rng.integers(0, 3, n).astype(np.int32)
    w = np.zeros((n, 12), np.float32)
    for i in range(n):
        a = rng.normal(size=(3, 3))
        q, _ = np.linalg.qr(a)
        m = q * rng.uniform(0.5, 2.0, 3)
        w[i, [0, 1, 2, 4, 5, 6, 8, 9, 10]] = m.reshape(-1)
        w[i, [3, 7, 11]] = rng.uniform(-10, 10, 3)
    ib, sb = oracle.instance_boxes(w, lod_boxes, il)
    b = oracle.lbvh_build_instances(w, lod_boxes, il, sb)
    order = np.argsort(b["sorted_keys"], kind="stable")
    assert np.array_equal(b["sorted_keys"][order], b["sorted_keys"])
    assert np.array_equal(b["parents"], LC.karras_bruteforce(b["sorted_keys"]))
    leaves = LC.walk_collapsed(b["nodes"], b["num_internal"], n)
    assert sum(c for _, c, _, _ in leaves) == n
    assert (sb[:3] <= ib[:, :3].min(axis=0)).all() and (sb[3:] >= ib[:, 3:].max(axis=0)).all()


@pytest.mark.parametrize("n,seed,mu", [(2, 1, -5.0), (1500, 2, -5.0), (60000, 3, -3.0), (200000, 4, -6.0)])
def test_parallel_build_equals_serial(oracle, n, seed, mu):
    """The OpenMP-parallel build (the CPU baseline of the reference's parallel
    path) returns the serial build's stages bit for bit, SAH costs included,
    also where counts saturate (LoDs of 200k triangles have nodes over 1023)."""
    v, t = S.random_triangle_soup(n, seed=seed, mu=mu)
    a = oracle.lbvh_build_tris(v, t)
    b = oracle.lbvh_build_tris_omp(v, t)
    assert a["num_internal"] == b["num_internal"]
    for k in ("sorted_keys", "sorted_ids", "parents", "offsets", "nodes"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["boxes"].view(np.uint32), b["boxes"].view(np.uint32))


def test_parallel_build_duplicate_keys(oracle):
    v = np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (4000, 1))
    t = np.arange(12000, dtype=np.uint32).reshape(-1, 3)
    a = oracle.lbvh_build_tris(v, t)
    b = oracle.lbvh_build_tris_omp(v, t)
    for k in ("sorted_keys", "sorted_ids", "parents", "offsets", "nodes"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["boxes"].view(np.uint32), b["boxes"].view(np.uint32))
