"""Closed-form pins of the oracle's path-tracing estimator (pt_sample,
pt_common.hpp:75-156): next-event estimation to point, area and background
lights, the MIS weights of NEE and of emitter hits (pt_common.hpp:115-118,
139-141), the light-tree choice probabilities (light_tree_sampling.hpp:195-413)
and the Russian-roulette division (random_walk.hpp:38-112).

None of these restate the oracle's code:
* a point light over a Lambert plane gives, sample by sample, exactly
  rho/pi * I cos(theta) / d^2 at the sample's hit point;
* a white-furnace plane under a uniform background converges to rho * Le;
* a Lambert plane under horizontal quad emitters (plus point lights)
  converges to rho/pi * E, with E from Lambert's polygon formula
  E = Le/2 * sum_i theta_i (n . Gamma_i), integrated over each pixel's
  footprint by Gauss-Legendre quadrature;
* the Cornell box agrees in expectation with an independent FP64 estimator
  (oracle/pt_fp64.cpp: its own BVH, light choice by area, no MIS, no
  roulette).
Each planted estimator mistake (oracle.set_mutation) must fail at least one
of these checks. Statistical checks use the oracle's own per-pixel sample
variance: |mean - expected| <= 5 standard errors (+ a quadrature allowance).
"""
import numpy as np
import pytest

import pt_util as P
from closed_form_util import (RHO, camera as _camera, expected_plane_radiance, oracle_scene as _oracle_scene,
                              plane_scene as _plane_scene, polygon_irradiance)
from mufflon_amd import hip

# planted mistakes of the oracle (pt_oracle.cpp, enum Mutation)
MUTATIONS = {
    1: "NEE to an area light without the cosine at the light",
    2: "NEE MIS denominator with 1 instead of the NEE count",
    3: "hit MIS without the NEE count",
    4: "light_pdf without the light-tree choice probability",
    5: "Russian roulette survivors not divided by the continuation probability",
    6: "point light without the 1/d^2 falloff",
    7: "background NEE without the division by the direction pdf",
    8: "NEE MIS hit pdf without the cosine at the light",
    9: "hit MIS with the incident pdf left in solid angle",
}


def _stat_check(O, ps, expected, iterations, nee, max_len, seed=1, rel_quad=2e-5):
    """Renders with the oracle; compares the aggregate over all pixels per
    channel with the expected image. Returns (ok, report)."""
    means, var, _ = O.pt_render_targets(ps, seed, 0, max_len, nee, iterations, [0], variance=[0])
    m = means[0].astype(np.float64)
    v = var[0].astype(np.float64)
    n = m.shape[0] * m.shape[1]
    agg, exp = m.sum((0, 1)) / n, expected.sum((0, 1)) / n
    se = np.sqrt(v.sum((0, 1)) / iterations) / n
    dev = np.abs(agg - exp)
    ok = bool((dev <= 5.0 * se + rel_quad * np.abs(exp)).all())
    return ok, "mean %s expected %s se %s (%.1f se)" % (agg, exp, se, float((dev / se).max()))


# ------------------------------------------------------------------ the checks
def _point_light_render(O, pl, nee, seed):
    """One iteration (the running means are the samples) of the plane under
    point lights pl; returns the radiance samples and, per light, rho/pi * I
    cos / (d^2 + 1e-10) at each sample's hit point (the Normal target is
    guide weight * (0, 1, 0), the Position target guide weight * position)."""
    cam = _camera(20)
    c, _ = _plane_scene()
    ps = _oracle_scene(O, c, cam, point_lights=pl)
    means, _ = O.pt_render_targets(ps, seed, 0, 3, nee, 1, [0, 2, 4])
    g = means[4][..., 1].astype(np.float64)
    x = means[2].astype(np.float64) / g[..., None]
    per_light = []
    for p in pl:
        d = p["position"].astype(np.float64) - x
        r2 = (d * d).sum(-1)
        L = (d[..., 1] / np.sqrt(r2) / (r2 + 1e-10))[..., None] * p["intensity"].astype(np.float64)
        per_light.append(L * np.asarray(RHO) / np.pi)
    return means[0].astype(np.float64), per_light


def check_point_light(O, nee=1, seed=1):
    """One point light, every sample: radiance == rho/pi * I cos / d^2."""
    pl = np.zeros(1, hip.POINT_LIGHT_DTYPE)
    pl["position"] = [(0.3, 0.8, -0.2)]
    pl["intensity"] = [(5.0, 4.0, 3.0)]
    val, (L,) = _point_light_render(O, pl, nee, seed)
    rel = np.abs(val - L) / L
    return bool(rel.max() < 2e-5), "max rel %.3g" % rel.max()


def check_two_point_lights(O, seed=1):
    """Two point lights, one NEE sample: every sample is L_j / p_j for the
    light j it chose, with choice probabilities p_j fixed over the image and
    p_0 + p_1 = 1, so the expectation is L_0 + L_1."""
    pl = np.zeros(2, hip.POINT_LIGHT_DTYPE)
    pl["position"] = [(0.3, 0.8, -0.2), (-0.4, 0.35, 0.5)]
    pl["intensity"] = [(5.0, 4.0, 3.0), (0.5, 1.5, 2.5)]
    val, Ls = _point_light_render(O, pl, 1, seed)
    ratios = [val / L for L in Ls]  # (h, w, 3) each
    # the chosen light: the one whose ratio is the same in all channels
    spread = [np.ptp(r, axis=-1) / r.mean(-1) for r in ratios]
    chose0 = spread[0] < spread[1]
    inv = []
    for j, mask in ((0, chose0), (1, ~chose0)):
        if not mask.any():
            return False, "light %d never chosen" % j
        r = ratios[j][mask]
        if np.ptp(r) > 1e-4 * r.mean():
            return False, "light %d: ratio not constant (%.3g .. %.3g)" % (j, r.min(), r.max())
        inv.append(1.0 / r.mean())
    total = sum(inv)
    return bool(abs(total - 1.0) < 1e-4), "choice probabilities %s sum %.6f" % (inv, total)


def check_furnace(O, nee=1, max_len=2, iterations=48):
    """White furnace: a Lambert plane under a uniform background Le converges
    to rho * Le (its hemisphere sees only the background)."""
    cam = _camera(16)
    c, _ = _plane_scene()
    bg = hip.background_record(hip.BACKGROUND_COLORED, color=(0.3, 0.4, 0.6), scale=0.8)
    ps = _oracle_scene(O, c, cam, background=bg)
    le = np.array([0.3, 0.4, 0.6]) * 0.8
    expected = np.broadcast_to(np.asarray(RHO) * le, (16, 16, 3))
    return _stat_check(O, ps, expected, iterations, nee, max_len)


AREA_EMITTERS = ((-0.6, -0.1, -0.5, 0.0, 0.6, (3.0, 2.0, 1.0)), (0.25, 0.45, -0.3, 0.1, 0.6, (1.0, 4.0, 6.0)))
# low and behind the camera: the plane sees it at grazing angles (small cosine at the light)
GRAZING_EMITTER = ((-1.0, 1.0, 1.6, 2.6, 0.4, (2.0, 2.0, 2.0)),)


def check_area_lights(O, nee=1, emitters=AREA_EMITTERS[:1], point_lights=None, pos_guide=False, iterations=256,
                      max_len=3):
    """Lambert plane under horizontal quad emitters (and point lights):
    converges to rho/pi * E with E from Lambert's polygon formula."""
    cam = _camera(16)
    c, polys = _plane_scene(emitters)
    ps = _oracle_scene(O, c, cam, point_lights=point_lights, pos_guide=pos_guide)
    expected = expected_plane_radiance(cam, polys, () if point_lights is None else point_lights)
    return _stat_check(O, ps, expected, iterations, nee, max_len)


def _mixed_point_lights():
    pl = np.zeros(2, hip.POINT_LIGHT_DTYPE)
    pl["position"] = [(0.5, 0.3, -0.8), (-0.9, 0.45, -0.2)]
    pl["intensity"] = [(0.4, 0.3, 0.2), (0.2, 0.2, 0.5)]
    return pl


def check_all(O):
    """Every closed-form check; returns {name: (ok, report)}."""
    return {
        "point nee1": check_point_light(O, 1),
        "point nee3": check_point_light(O, 3, seed=2),
        "two points": check_two_point_lights(O),
        "furnace nee1": check_furnace(O, 1),
        "furnace nee3 depth5": check_furnace(O, 3, max_len=5),
        "area nee1": check_area_lights(O, 1),
        "area nee2": check_area_lights(O, 2),
        "grazing nee1": check_area_lights(O, 1, GRAZING_EMITTER),
        "mixed nee2": check_area_lights(O, 2, AREA_EMITTERS, _mixed_point_lights()),
        "mixed nee2 guide": check_area_lights(O, 2, AREA_EMITTERS, _mixed_point_lights(), pos_guide=True),
    }


# ------------------------------------------------------------------ tests
def test_polygon_formula_against_quadrature():
    """The closed form itself: Lambert's polygon formula vs a fine midpoint
    quadrature of cos*cos/r^2 over the quad."""
    poly = np.array([[-0.6, 0.6, -0.5], [-0.1, 0.6, -0.5], [-0.1, 0.6, 0.0], [-0.6, 0.6, 0.0]])
    x = np.array([[0.2, 0.0, 0.3], [-0.35, 0.0, -0.25], [1.5, 0.0, -2.0]])
    n = 800
    s = (np.arange(n) + 0.5) / n
    gx, gz = np.meshgrid(-0.6 + 0.5 * s, -0.5 + 0.5 * s, indexing="ij")
    pts = np.stack([gx.ravel(), np.full(gx.size, 0.6), gz.ravel()], 1)
    for xi, e in zip(x, polygon_irradiance(x, poly)):
        d = pts - xi
        r2 = (d * d).sum(1)
        quad = ((d[:, 1] ** 2) / r2 ** 2).mean() * 0.25
        assert abs(e - quad) < 1e-5 * quad


@pytest.mark.parametrize("nee", [1, 3])
def test_point_light_exact_per_sample(oracle, nee):
    ok, rep = check_point_light(oracle, nee, seed=nee)
    assert ok, rep


def test_two_point_lights_choice(oracle):
    ok, rep = check_two_point_lights(oracle)
    assert ok, rep


@pytest.mark.parametrize("nee,max_len", [(1, 2), (3, 5)])
def test_white_furnace(oracle, nee, max_len):
    ok, rep = check_furnace(oracle, nee, max_len)
    assert ok, rep


@pytest.mark.parametrize("nee", [1, 2])
def test_quad_light_irradiance(oracle, nee):
    ok, rep = check_area_lights(oracle, nee)
    assert ok, rep


def test_grazing_quad_light(oracle):
    ok, rep = check_area_lights(oracle, 1, GRAZING_EMITTER)
    assert ok, rep


@pytest.mark.parametrize("guide", [False, True])
def test_two_quads_and_point_lights(oracle, guide):
    ok, rep = check_area_lights(oracle, 2, AREA_EMITTERS, _mixed_point_lights(), pos_guide=guide)
    assert ok, rep


def test_cornell_matches_fp64_estimator(oracle):
    """The oracle's Cornell box (NEE through the light tree, MIS, roulette)
    and the independent FP64 estimator converge to the same image: per-pixel
    z-scores have mean 0 and mean square 1, the aggregate agrees within 5 SE."""
    c = P.cornell(8, 16, 16)
    ps = P.oracle_pt_scene(oracle, c)
    its = 512
    means, var, _ = oracle.pt_render_targets(ps, 1, 0, 8, 1, its, [0], variance=[0])
    mo, vo = means[0].astype(np.float64), var[0].astype(np.float64)
    m64, v64 = oracle.fp64_render(c["verts"], c["normals"], c["tris"], c["mat_idx"], c["materials"], c["camera"], 8,
                                  4096, seed=3)
    se2 = vo / its + v64 / 4096
    ok = se2 > 0
    z = (mo - m64)[ok] / np.sqrt(se2[ok])
    n = ok.sum()
    assert abs(z.mean()) < 5.0 / np.sqrt(n), z.mean()
    assert abs((z * z).mean() - 1.0) < 5.0 * np.sqrt(2.0 / n) + 0.05, (z * z).mean()
    agg_se = np.sqrt(se2.sum((0, 1)))
    assert (np.abs(mo.sum((0, 1)) - m64.sum((0, 1))) < 5.0 * agg_se).all()


def test_fp64_estimator_closed_forms(oracle):
    """The FP64 estimator itself on the closed forms: furnace and quad light."""
    cam = _camera(12)
    c, polys = _plane_scene(AREA_EMITTERS)
    pl = _mixed_point_lights()
    m, v = oracle.fp64_render(c["verts"], c["normals"], c["tris"], c["mat_idx"], c["materials"], cam, 3, 2048,
                              point_lights=pl)
    exp = expected_plane_radiance(cam, polys, pl)
    se = np.sqrt(v.sum((0, 1)) / 2048)
    assert (np.abs(m.sum((0, 1)) - exp.sum((0, 1))) < 5 * se + 2e-5 * exp.sum((0, 1))).all()
    c, _ = _plane_scene()
    le = np.array([0.24, 0.32, 0.48])
    m, v = oracle.fp64_render(c["verts"], c["normals"], c["tris"], c["mat_idx"], c["materials"], cam, 2, 2048,
                              background=le)
    se = np.sqrt(v.sum((0, 1)) / 2048)
    exp = np.asarray(RHO) * le * 144  # float32 background: 1e-6 relative allowance
    assert (np.abs(m.sum((0, 1)) - exp) < 5 * se + 1e-6 * exp).all()


@pytest.mark.parametrize("mutation", sorted(MUTATIONS))
def test_planted_mistake_is_caught(oracle, mutation):
    """Each planted estimator mistake fails at least one closed-form check."""
    old = oracle.set_mutation(mutation)
    try:
        results = check_all(oracle)
    finally:
        oracle.set_mutation(old)
    failed = [k for k, (ok, _) in results.items() if not ok]
    assert failed, "mutation %d (%s) passed every check: %s" % (mutation, MUTATIONS[mutation], results)
