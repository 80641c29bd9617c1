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
