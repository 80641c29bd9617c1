"""GPU closest-hit / any-hit (through the C-ABI) against the oracle tracer on
identical rays: hit instance and primitive ids bit-exact, t and barycentrics
bit-exact (both sides evaluate the same IEEE operations in the same order)."""
import numpy as np
import pytest

import scene_util as U
from mufflon_amd import scenes as S

pytestmark = pytest.mark.gpu


def _gpu_scene(hip_ctx, geoms, inst_lod, w):
    lods = []
    for v, t in geoms:
        lid = hip_ctx.lod_create(v, t)
        hip_ctx.lbvh_build(lid)
        lods.append(lid)
    return hip_ctx.scene_create(np.array(lods, np.int32)[inst_lod], w)


def _compare(hip_ctx, oracle, sc, oscene, rays):
    print("trace parity rays=%d" % len(rays), flush=True)
    gt, gi, gp, gb, st = hip_ctx.trace_closest(sc, rays)
    ot, oi, op, obary = oracle.trace_closest(oscene, rays)
    assert (st == 0).all(), "traversal stack overflow"
    assert np.array_equal(gi, oi), "instance ids"
    assert np.array_equal(gp, op), "primitive ids"
    assert np.array_equal(gt.view(np.uint32), ot.view(np.uint32)), "t"
    hit = oi >= 0
    assert np.array_equal(gb[hit].view(np.uint32), obary[hit].view(np.uint32)), "barycentrics"
    go = hip_ctx.trace_any(sc, rays)
    oo = oracle.trace_any(oscene, rays)
    assert np.array_equal(go, oo), "any hit"
    return hit


def test_single_instance_trace(hip_ctx, oracle):
    v, t = S.random_triangle_soup(100_000, seed=21, mu=-4.5)
    w = S.IDENTITY34.reshape(1, 12)
    sc = _gpu_scene(hip_ctx, [(v, t)], np.array([0], np.int32), w)
    oscene, _ = U.oracle_scene(oracle, [(v, t)], np.array([0], np.int32), w)
    rays = np.concatenate([S.pinhole_rays(256, 128), S.surface_rays(v, t, 30000, seed=21)])
    hit = _compare(hip_ctx, oracle, sc, oscene, rays)
    assert hit.sum() > 1000


def test_multi_instance_trace(hip_ctx, oracle):
    geoms, rng = U.small_lods(11)
    n_inst = 300
    il = rng.integers(0, len(geoms), n_inst).astype(np.int32)
    w = U.random_transforms(n_inst, 12)
    sc = _gpu_scene(hip_ctx, geoms, il, w)
    oscene, _ = U.oracle_scene(oracle, geoms, il, w)
    rays = U.rays_toward(rng, 50_000)
    hit = _compare(hip_ctx, oracle, sc, oscene, rays)
    assert hit.sum() > 1000


def test_empty_and_short_rays(hip_ctx, oracle):
    v, t = S.random_triangle_soup(1000, seed=2, mu=-3.0)
    w = S.IDENTITY34.reshape(1, 12)
    sc = _gpu_scene(hip_ctx, [(v, t)], np.array([0], np.int32), w)
    oscene, _ = U.oracle_scene(oracle, [(v, t)], np.array([0], np.int32), w)
    rays = S.pinhole_rays(32, 32)
    rays[:, 3] = 1e-3  # shorter than any distance to the soup
    _compare(hip_ctx, oracle, sc, oscene, rays)
    empty = np.zeros((0, 8), np.float32)
    g = hip_ctx.trace_closest(sc, empty)
    assert len(g[0]) == 0


@pytest.mark.slow
def test_lbvh_trace_workload_full(hip_ctx, oracle):
    """1M random triangles, 2M coherent + 2M incoherent rays."""
    v, t, rays = S.lbvh_trace_workload()
    w = S.IDENTITY34.reshape(1, 12)
    sc = _gpu_scene(hip_ctx, [(v, t)], np.array([0], np.int32), w)
    oscene, _ = U.oracle_scene(oracle, [(v, t)], np.array([0], np.int32), w)
    hit = _compare(hip_ctx, oracle, sc, oscene, rays)
    assert hit.sum() > 100_000


def test_visit_counts_match_oracle(hip_ctx, oracle):
    """The GPU traversal visits the same nodes and triangles as the oracle."""
    v, t = S.random_triangle_soup(50_000, seed=31, mu=-4.5)
    w = S.IDENTITY34.reshape(1, 12)
    sc = _gpu_scene(hip_ctx, [(v, t)], np.array([0], np.int32), w)
    oscene, _ = U.oracle_scene(oracle, [(v, t)], np.array([0], np.int32), w)
    rays = np.concatenate([S.pinhole_rays(128, 64), S.surface_rays(v, t, 8000, seed=31)])
    gn, gp = hip_ctx.trace_stats(sc, rays)
    _, _, _, _, on, op = oracle.trace_closest(oscene, rays, counts=True)
    assert np.array_equal(gn, on) and np.array_equal(gp, op)


def _staircase(n=24):
    """Axis-aligned geometry with coordinates on a 1/8 grid: the top faces
    and the two visible side faces of n x n unit columns of random integer
    heights (2 triangles per face), so that box faces, triangle edges and
    shared vertices coincide exactly."""
    rng = np.random.default_rng(5)
    h = rng.integers(1, 6, (n, n)).astype(np.float32) / 8.0
    verts, tris = [], []

    def quad(a, b, c, d):
        k = len(verts)
        verts.extend([a, b, c, d])
        tris.extend([(k, k + 1, k + 2), (k, k + 2, k + 3)])

    for i in range(n):
        for j in range(n):
            x0, x1, z0, z1, y = i / 8.0, (i + 1) / 8.0, j / 8.0, (j + 1) / 8.0, h[i, j]
            quad((x0, y, z0), (x1, y, z0), (x1, y, z1), (x0, y, z1))
            quad((x0, 0, z0), (x1, 0, z0), (x1, y, z0), (x0, y, z0))
            quad((x0, 0, z0), (x0, y, z0), (x0, y, z1), (x0, 0, z1))
    return np.array(verts, np.float32), np.array(tris, np.uint32)


def _adversarial_rays(rng, n, extent):
    """Rays whose direction has exactly zero components (their slab tests see
    infinities and NaN), tiny components (|1/d| up to 1e30), rays in the
    planes of faces and along shared edges, and ordinary ones, shuffled so
    that waves mix every kind."""
    parts = []
    # axis-aligned: origins on the 1/16 grid, directions +-x, +-y, +-z
    o = rng.integers(-4, int(extent * 16) + 4, (n, 3)).astype(np.float32) / 16.0
    d = np.zeros((n, 3), np.float32)
    ax = rng.integers(0, 3, n)
    d[np.arange(n), ax] = rng.choice([-1.0, 1.0], n)
    parts.append((o, d))
    # one or two tiny components
    o = rng.uniform(-0.5, extent + 0.5, (n, 3)).astype(np.float32)
    d = np.zeros((n, 3), np.float32)
    ax = rng.integers(0, 3, n)
    d[np.arange(n), ax] = rng.choice([-1.0, 1.0], n)
    d += (rng.choice([-1, 1], (n, 3)) * 10.0 ** rng.uniform(-30, -3, (n, 3))).astype(np.float32) * (d == 0)
    parts.append((o, d))
    # in a face plane (y on the height grid) and along grid lines (x, z on the grid)
    o = np.stack([rng.integers(0, int(extent * 8), n) / 8.0, rng.integers(1, 6, n) / 8.0,
                  rng.uniform(-1, extent + 1, n)], -1).astype(np.float32)
    d = np.stack([rng.uniform(-1, 1, n), np.zeros(n), rng.uniform(-1, 1, n)], -1).astype(np.float32)
    parts.append((o, d))
    o = np.stack([rng.uniform(-1, extent + 1, n), np.full(n, 3.0), rng.uniform(-1, extent + 1, n)], -1)
    d = np.stack([rng.uniform(-0.3, 0.3, n), -np.ones(n), rng.uniform(-0.3, 0.3, n)], -1)
    parts.append((o.astype(np.float32), d.astype(np.float32)))
    o = np.concatenate([p[0] for p in parts])
    d = np.concatenate([p[1] for p in parts]).astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((len(o), 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = S.MAX_SCENE_SIZE
    rays[:, 4:7] = d
    return rays[rng.permutation(len(rays))]


@pytest.mark.parametrize("multi", [False, True])
def test_adversarial_rays(hip_ctx, oracle, multi):
    """Rays whose slab tests see infinities and NaN (zero direction
    components) next to ordinary and nearly axis-parallel ones in the same
    waves, on geometry whose faces, edges and boxes coincide: hits, any-hit
    and the visit counts of the reference order equal the oracle's."""
    v, t = _staircase()
    rng = np.random.default_rng(17)
    if multi:
        n_inst = 40
        il = np.zeros(n_inst, np.int32)
        w = np.zeros((n_inst, 12), np.float32)
        w[:, [0, 5, 10]] = 1.0
        w[:, 3] = -(np.arange(n_inst) % 8) * 3.0  # translations only: object rays keep their zero components
        w[:, 11] = -(np.arange(n_inst) // 8) * 3.0
    else:
        il = np.array([0], np.int32)
        w = S.IDENTITY34.reshape(1, 12)
    sc = _gpu_scene(hip_ctx, [(v, t)], il, w)
    oscene, _ = U.oracle_scene(oracle, [(v, t)], il, w)
    rays = _adversarial_rays(rng, 6000, 24 / 8.0 * (8 if multi else 1))
    print("adversarial rays: %d, multi-instance %s" % (len(rays), multi), flush=True)
    hit = _compare(hip_ctx, oracle, sc, oscene, rays)
    assert hit.sum() > 1000
    n_g, t_g = hip_ctx.trace_stats(sc, rays)
    _, _, _, _, n_o, t_o = oracle.trace_closest(oscene, rays, counts=True)
    assert np.array_equal(n_g, n_o) and np.array_equal(t_g, t_o), "visit counts"
