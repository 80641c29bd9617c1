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
* a faceted Lambert sphere with smooth normals and a rough glass sphere in
  front of an emitting wall: per object (and, on the faceted sphere, pixel
  by pixel) the light tracer agrees with the path tracer, which is where the
  adjoint shading normal correction and the eta^2 of adjoint refraction show;
* determinism and the Density / Lightness targets' basic properties.
Statistical checks compare image aggregates within 5 standard errors from
the recorded per-pixel variances. Each planted light-tracer mistake of the
oracle (oracle.set_mutation 10-14) must fail at least one of these checks."""
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


LT_MUTATIONS = {
    10: "emission not divided by the light's choice probability",
    11: "adjoint BSDF without the shading normal correction",
    12: "adjoint refraction without eta^2",
    13: "camera importance without 1 / cos^3",
    14: "area light seen by the camera without its cosine",
}


def _spheres_scene(oracle):
    """A coarse Lambert sphere with smooth normals (4 x 6 facets: the adjoint
    shading normal correction is far from 1) and a rough glass sphere (Walter,
    ior 1.5: light enters it on sampled refractions and leaves towards the
    camera on evaluated ones, so eta^2 does not cancel) in front of an emitting
    wall, a point light and a floor; returns the oracle scene and the triangle
    ranges of the two spheres."""
    mb = S.MeshBuilder()
    mb.add_quad((-2, 0, 1.5), (2, 0, 1.5), (2, 0, -1.5), (-2, 0, -1.5), (0, 1, 0), 0)
    n0 = len(mb.arrays()[3])
    mb.add_uv_sphere((-0.45, 0.45, 0.0), 0.38, 4, 6, 0)
    n1 = len(mb.arrays()[3])
    mb.add_uv_sphere((0.45, 0.45, 0.0), 0.38, 24, 48, 1)
    n2 = len(mb.arrays()[3])
    mb.add_quad((-2, 0, -1), (2, 0, -1), (2, 2, -1), (-2, 2, -1), (0, 0, 1), 2)
    v, n, uv, t, m = mb.arrays()
    mats = np.concatenate([S.material(S.MAT_LAMBERT, albedo=(0.8, 0.7, 0.6)),
                           S.material(S.MAT_WALTER, roughness=0.3, ior=1.5, absorption=(0.0, 0.0, 0.0)),
                           S.material(S.MAT_EMISSIVE, radiance=(2.0, 2.0, 2.0), scale=1.0)])
    cam = dict(position=np.array([0.0, 0.45, 2.2], np.float32), view_dir=np.array([0, 0, -1], np.float32),
               up=np.array([0, 1, 0], np.float32), tan_vfov=np.float32(np.tan(np.radians(15.0))), width=24, height=16)
    c = dict(verts=v, normals=n, uvs=uv, tris=t, mat_idx=m, materials=mats)
    ps = _oracle_scene(oracle, c, cam, point_lights=_point((-0.2, 1.4, 1.2), (1.5, 1.5, 1.5)))
    return ps, (n0, n1), (n1, n2)


def _lt_checks(oracle):
    """The light tracer pins above at reduced sample counts; (name, passed)."""
    out = []
    # per-object agreement with the path tracer on the pixels whose primary
    # hit is the coarse sphere, resp. the glass sphere
    ps, coarse, glass = _spheres_scene(oracle)
    n_lt, n_pt = 3000, 1500
    mean, var = oracle.lt_render(ps, 5, 0, 6, n_lt, variance=True)
    pm, pv, _ = oracle.pt_render_targets(ps, 7, 0, 6, 1, n_pt, [0, hip.TARGET_HIT_ID], variance=[0])
    hit = pm[hip.TARGET_HIT_ID][..., 1]
    for label, (a, b) in (("coarse sphere", coarse), ("glass sphere", glass)):
        sel = (hit >= a) & (hit < b)  # the mean id of the pixel's primary hits
        k = int(sel.sum())
        m_lt = mean[..., :3][sel].astype(np.float64).sum(0) / k
        se_lt = np.sqrt(var[..., :3][sel].astype(np.float64).sum(0) / n_lt) / k
        m_pt = pm[0][sel].astype(np.float64).sum(0) / k
        se_pt = np.sqrt(pv[0][sel].astype(np.float64).sum(0) / n_pt) / k
        out.append((label, bool((np.abs(m_lt - m_pt) <= 5 * np.sqrt(se_lt ** 2 + se_pt ** 2)).all())))
        if label == "coarse sphere":
            # pixel by pixel: the shading normal correction changes the
            # pixels of a faceted sphere in both directions
            z = (mean[..., 0][sel] - pm[0][..., 0][sel]) / np.sqrt(var[..., 0][sel] / n_lt + pv[0][..., 0][sel] / n_pt)
            out.append(("coarse sphere pixels", bool(float((z.astype(np.float64) ** 2).mean()) < 2.0)))
    cam = _camera(16)
    c, _ = _plane_scene()
    pl = np.concatenate([_point((0.2, 0.9, 0.1), (2.0, 1.5, 1.0)), _point((-0.5, 0.7, -0.2), (1.0, 1.0, 1.0))])
    ps = _oracle_scene(oracle, c, cam, point_lights=pl)
    mean, var = oracle.lt_render(ps, 1, 2, 2, 800, variance=True)
    m, se = _aggregate(mean[..., :3], var[..., :3], 800)
    exp = expected_plane_radiance(cam, [], pl).sum((0, 1)) / (16 * 16)
    out.append(("point lights", bool((np.abs(m - exp) <= 5 * se + 2e-5 * exp).all())))
    for label, ps2, lo, hi in (("cornell", P.oracle_pt_scene(oracle, P.cornell(8, 16, 24)), 0, 5),
                               ("mixed room", _mixed_room(oracle), 2, 4)):
        mean, var = oracle.lt_render(ps2, 5, lo, hi, 600, variance=True)
        m_lt, se_lt = _aggregate(mean[..., :3], var[..., :3], 600)
        pm, pv, _ = oracle.pt_render_targets(ps2, 7, lo, hi, 1, 300, [0], variance=[0])
        m_pt, se_pt = _aggregate(pm[0], pv[0], 300)
        out.append((label, bool((np.abs(m_lt - m_pt) <= 5 * np.sqrt(se_lt ** 2 + se_pt ** 2)).all())))
    return out


def _mixed_room(oracle):
    sc = S.mixed_scene(24, 20, instanced=True)
    area = np.concatenate([S.emissive_area_lights_ex(sc["lods"][l], sc["materials"], i, sc["world_to_inst"][i])
                           for i, l in enumerate(sc["inst_lod"])])
    geoms = [(L["verts"], L["tris"], L.get("quads"), L.get("spheres")) for L in sc["lods"]]
    _, lods = U.oracle_scene(oracle, geoms, sc["inst_lod"], sc["world_to_inst"])
    _, box = oracle.instance_boxes(sc["world_to_inst"], np.stack([L["box"] for L in lods]), sc["inst_lod"])
    box = np.asarray(box, np.float32)
    tree = hip.light_tree_build(area, sc["point_lights"], box, sc["materials"])
    return P.oracle_pt_scene_general(oracle, sc, tree, box)


def test_light_tracer_checks_pass_unmutated(oracle):
    res = _lt_checks(oracle)
    print(res)
    assert all(ok for _, ok in res), res


@pytest.mark.parametrize("mutation", sorted(LT_MUTATIONS))
def test_planted_light_tracer_mistake_is_caught(oracle, mutation):
    old = oracle.set_mutation(mutation)
    try:
        res = _lt_checks(oracle)
    finally:
        oracle.set_mutation(old)
    print(LT_MUTATIONS[mutation], res)
    assert not all(ok for _, ok in res), "planted mistake not caught: " + LT_MUTATIONS[mutation]
