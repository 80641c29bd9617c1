"""The oracle's light tracer (lt_sample, lt_common.hpp:17-78) against what
the mathematics fixes, not against a restatement of its own code:
* a point light over a Lambert plane: every pixel converges to
  rho/pi * I cos(theta) / d^2 averaged over the pixel's footprint (the
  camera's importance makes the light tracer estimate the same pixel
  integral the path tracer does);
* horizontal area emitters (triangles) over the plane: Lambert's polygon
  formula;
* the Cornell box, whose smooth sphere normals make the adjoint shading
  normal correction matter, and a room of quads, spheres and a glass
  sphere: the light tracer agrees in expectation with the path tracer
  (itself pinned by closed forms and an independent FP64 estimator) over
  the same path lengths;
* determinism and the Density / Lightness targets' basic properties.
Statistical checks compare image aggregates within 5 standard errors from
the recorded per-pixel variances."""
import numpy as np
import pytest

import pt_util as P
import scene_util as U
from closed_form_util import (RHO, camera as _camera, expected_plane_radiance, oracle_scene as _oracle_scene,
                              plane_scene as _plane_scene)
from mufflon_amd import hip
from mufflon_amd import scenes as S


def _point(pos, intensity):
    p = np.zeros(1, hip.POINT_LIGHT_DTYPE)
    p["position"] = [pos]
    p["intensity"] = [intensity]
    return p


def _aggregate(mean, var, iterations):
    n = mean.shape[0] * mean.shape[1]
    m = mean.astype(np.float64).sum((0, 1)) / n
    se = np.sqrt(var.astype(np.float64).sum((0, 1)) / iterations) / n
    return m, se


def test_point_light_over_plane(oracle):
    cam = _camera(16)
    c, _ = _plane_scene()
    pl = _point((0.2, 0.9, 0.1), (2.0, 1.5, 1.0))
    ps = _oracle_scene(oracle, c, cam, point_lights=pl)
    its = 3000
    mean, var = oracle.lt_render(ps, 1, 2, 2, its, variance=True)
    m, se = _aggregate(mean[..., :3], var[..., :3], its)
    exp = expected_plane_radiance(cam, [], pl).sum((0, 1)) / (16 * 16)
    dev = np.abs(m - exp)
    print("lt point light: mean %s expected %s (%.1f se)" % (m, exp, float((dev / se).max())))
    assert (dev <= 5 * se + 2e-5 * exp).all()
    # the pixels agree one by one too (with a looser bound for the many tests)
    e = expected_plane_radiance(cam, [], pl)
    z = np.abs(mean[..., :3] - e) / np.sqrt(var[..., :3] / its + 1e-20)
    assert np.mean(z < 4.0) > 0.99


def test_area_emitters_over_plane(oracle):
    cam = _camera(16)
    c, polys = _plane_scene(emitters=[(-0.4, 0.2, -0.3, 0.4, 0.8, (3.0, 2.5, 2.0)),
                                      (0.3, 0.7, 0.0, 0.5, 0.6, (1.0, 2.0, 4.0))])
    ps = _oracle_scene(oracle, c, cam)
    its = 3000
    mean, var = oracle.lt_render(ps, 3, 2, 2, its, variance=True)
    m, se = _aggregate(mean[..., :3], var[..., :3], its)
    exp = expected_plane_radiance(cam, polys).sum((0, 1)) / (16 * 16)
    dev = np.abs(m - exp)
    print("lt area emitters: mean %s expected %s (%.1f se)" % (m, exp, float((dev / se).max())))
    assert (dev <= 5 * se + 2e-5 * exp).all()


def _lt_vs_pt(oracle, ps, min_len, max_len, its_lt, its_pt, label):
    mean, var = oracle.lt_render(ps, 5, min_len, max_len, its_lt, variance=True)
    m_lt, se_lt = _aggregate(mean[..., :3], var[..., :3], its_lt)
    pm, pv, _ = oracle.pt_render_targets(ps, 7, min_len, max_len, 1, its_pt, [0], variance=[0])
    m_pt, se_pt = _aggregate(pm[0], pv[0], its_pt)
    dev = np.abs(m_lt - m_pt)
    se = np.sqrt(se_lt ** 2 + se_pt ** 2)
    print("%s: light tracer %s path tracer %s (%.1f se)" % (label, m_lt, m_pt, float((dev / se).max())))
    assert (dev <= 5 * se).all()
    assert m_pt.min() > 0


def test_agrees_with_path_tracer_cornell(oracle):
    c = P.cornell(8, 16, 24)
    ps = P.oracle_pt_scene(oracle, c)
    _lt_vs_pt(oracle, ps, 2, 5, 1500, 600, "cornell")


def test_agrees_with_path_tracer_mixed_room(oracle):
    sc = S.mixed_scene(24, 20, instanced=True)
    # the light tree of every instance's emissive triangles, quads and spheres
    area = np.concatenate([S.emissive_area_lights_ex(sc["lods"][l], sc["materials"], i, sc["world_to_inst"][i])
                           for i, l in enumerate(sc["inst_lod"])])
    geoms = [(L["verts"], L["tris"], L.get("quads"), L.get("spheres")) for L in sc["lods"]]
    _, lods = U.oracle_scene(oracle, geoms, sc["inst_lod"], sc["world_to_inst"])
    _, box = oracle.instance_boxes(sc["world_to_inst"], np.stack([L["box"] for L in lods]), sc["inst_lod"])
    box = np.asarray(box, np.float32)
    tree = hip.light_tree_build(area, sc["point_lights"], box, sc["materials"])
    ps = P.oracle_pt_scene_general(oracle, sc, tree, box)
    _lt_vs_pt(oracle, ps, 2, 4, 1500, 600, "mixed room")


def test_deterministic_and_auxiliary_targets(oracle):
    c = P.cornell(8, 16, 16)
    ps = P.oracle_pt_scene(oracle, c)
    a, _ = oracle.lt_render(ps, 2, 0, 4, 3, density=True)
    b, _ = oracle.lt_render(ps, 2, 0, 4, 3, density=True)
    assert np.array_equal(a, b)
    d, _ = oracle.lt_render(ps, 2, 0, 4, 3, density=False)
    # the Density target is recorded on request only; Radiance and Lightness
    # do not depend on it (the random decisions are the same)
    assert (d[..., 4] == 0).all() and (a[..., 4] > 0).any()
    assert np.array_equal(a[..., :4], d[..., :4])
    assert (a >= 0).all() and (a[..., 3] > 0).any()
    # another seed gives another image
    e, _ = oracle.lt_render(ps, 3, 0, 4, 3, density=True)
    assert not np.array_equal(a, e)
