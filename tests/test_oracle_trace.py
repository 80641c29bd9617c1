"""Pins the traversal oracle (oracle/trace_oracle.cpp) against an all-pairs
brute-force tracer on single- and multi-instance scenes, and checks any-hit
against closest-hit and the equal-distance tie rule."""
import numpy as np

import scene_util as U
from mufflon_amd import scenes as S


def test_single_instance_matches_bruteforce(oracle):
    v, t = S.random_triangle_soup(20000, seed=3, mu=-4.0)
    scene, _ = U.oracle_scene(oracle, [(v, t)], np.array([0], np.int32), S.IDENTITY34.reshape(1, 12))
    rays = np.concatenate([S.pinhole_rays(96, 48), S.surface_rays(v, t, 3000, seed=3)])
    tt, ii, pp, _ = oracle.trace_closest(scene, rays)
    bt, bi, bp = oracle.trace_brute(scene, rays)
    assert (ii >= 0).sum() > 500
    assert np.array_equal(ii, bi) and np.array_equal(pp, bp)
    assert np.array_equal(tt, bt)
    occ = oracle.trace_any(scene, rays)
    assert np.array_equal(occ == 1, ii >= 0)


def test_multi_instance_matches_bruteforce(oracle):
    geoms, rng = U.small_lods(7)
    n_inst = 40
    il = rng.integers(0, len(geoms), n_inst).astype(np.int32)
    w = U.random_transforms(n_inst, 8)
    scene, _ = U.oracle_scene(oracle, geoms, il, w)
    rays = U.rays_toward(rng, 4000)
    tt, ii, pp, _ = oracle.trace_closest(scene, rays)
    bt, bi, bp = oracle.trace_brute(scene, rays)
    assert (ii >= 0).sum() > 200
    assert np.array_equal(ii, bi) and np.array_equal(pp, bp)
    hit = ii >= 0
    # t goes through object-space scalings in a different order: ulp-level
    assert np.allclose(tt[hit], bt[hit], rtol=1e-6, atol=0)
    occ = oracle.trace_any(scene, rays)
    assert np.array_equal(occ == 1, hit)


def test_tmax_limits_hits(oracle):
    v, t = S.random_triangle_soup(5000, seed=4, mu=-3.5)
    scene, _ = U.oracle_scene(oracle, [(v, t)], np.array([0], np.int32), S.IDENTITY34.reshape(1, 12))
    rays = S.pinhole_rays(64, 64)
    tt, ii, _, _ = oracle.trace_closest(scene, rays)
    hit = ii >= 0
    short = rays.copy()
    short[:, 3] = np.where(hit, tt * np.float32(0.999), np.float32(1.0))
    ts, i2, _, _ = oracle.trace_closest(scene, short)
    # nothing can be hit strictly before the first hit
    assert (i2[hit] == -1).all()
    assert (oracle.trace_any(scene, short)[hit] == 0).all()


def test_equal_distance_tie_takes_lowest_id(oracle):
    # the same triangle three times: every hit must report the lowest id
    v, t = S.random_triangle_soup(200, seed=5, mu=-2.0)
    t3 = np.concatenate([t, t[::-1], t])
    scene, _ = U.oracle_scene(oracle, [(v, t3)], np.array([0], np.int32), S.IDENTITY34.reshape(1, 12))
    rays = S.pinhole_rays(128, 128)
    _, ii, pp, _ = oracle.trace_closest(scene, rays)
    hit = ii >= 0
    assert hit.sum() > 50
    assert (pp[hit] < 200).all()
    _, _, bp = oracle.trace_brute(scene, rays)
    assert np.array_equal(pp, bp)
