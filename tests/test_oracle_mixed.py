"""Quads and spheres in the oracle: the primitive tests of intersects_primitve
(intersection.cu:86-316: the bilinear patch of Ramsey et al., the ray/sphere
test, the quadratic), the mixed-primitive LBVH (centroids, boxes and the
three-type count code of build_lbvh.cu:285-317), the traversal over mixed
scenes, the quad and sphere light records of the light tree (lights.hpp:
108-137, lights.cpp:43-80) and their sampling (light_sampling.hpp:142-171,
376-432).

The pins do not restate the oracle: roots against numpy in double, patch hits
against the patch point they were aimed at, sphere hits against the distance
to the center, LBVH leaves against the primitive boxes and the per-node
collapse rule, traversal against the brute-force tracer, fluxes and pdfs
against their closed forms, and the path tracer against irradiance closed
forms (a sphere light acts as a point source of intensity pi r^2 L; a planar
quad light obeys Lambert's polygon formula whatever its parametrization)."""
import ctypes as C

import numpy as np
import pytest

import lbvh_check as LC
import scene_util as U
from closed_form_util import (RHO, camera as _camera, expected_plane_radiance, pixel_points, GAUSS3)
from mufflon_amd import hip
from mufflon_amd import scenes as S

f32p = C.POINTER(C.c_float)


def _lib(O):
    lib = O.load()
    if not hasattr(lib, "_mixed_ready"):
        lib.orc_ray_quad.argtypes = [f32p, f32p, f32p]
        lib.orc_ray_sphere.argtypes = [f32p, f32p, f32p]
        lib.orc_solve_square_poly.argtypes = [C.c_float, C.c_float, C.c_float, f32p]
        lib.orc_perpendicular.argtypes = [f32p, f32p]
        lib.orc_perpendicular.restype = None
        lib.orc_bilerp.argtypes = [f32p, C.c_float, C.c_float]
        lib.orc_bilerp.restype = C.c_float
        lib._mixed_ready = True
    return lib


def _f(a):
    a = np.ascontiguousarray(a, np.float32)
    return a, a.ctypes.data_as(f32p)


def ray_quad(O, o, d, quad):
    r, rp = _f(np.concatenate([o, d]))
    q, qp = _f(np.asarray(quad).reshape(-1))
    out, op = _f(np.zeros(3))
    hit = _lib(O).orc_ray_quad(rp, qp, op)
    return bool(hit), out


def ray_sphere(O, o, d, sphere):
    r, rp = _f(np.concatenate([o, d]))
    s, sp = _f(sphere)
    out, op = _f(np.zeros(1))
    hit = _lib(O).orc_ray_sphere(rp, sp, op)
    return bool(hit), float(out[0])


# ------------------------------------------------------------------ primitives
def test_solve_square_poly(oracle):
    lib = _lib(oracle)
    rng = np.random.default_rng(1)
    for _ in range(2000):
        a, b, c = rng.normal(size=3).astype(np.float32)
        out, op = _f(np.zeros(2))
        ok = lib.orc_solve_square_poly(a, b, c, op)
        disc = float(b) * float(b) - 4.0 * float(a) * float(c)
        if disc < -1e-6 * (float(b) ** 2 + abs(4.0 * a * c)):
            assert not ok
            continue
        if disc <= 1e-6 * (float(b) ** 2 + abs(4.0 * a * c)):
            continue  # rounding decides near a double root
        assert ok and out[0] <= out[1]
        r = np.sort(np.roots([float(a), float(b), float(c)]).real)
        np.testing.assert_allclose(out, r, rtol=2e-5, atol=1e-6 * max(1.0, abs(r).max()))
    out, op = _f(np.zeros(2))
    assert lib.orc_solve_square_poly(0.0, 2.0, -3.0, op) and out[0] == out[1] == np.float32(1.5)  # linear
    assert not lib.orc_solve_square_poly(0.0, 0.0, 1.0, op)
    assert lib.orc_solve_square_poly(1.0, 0.0, 0.0, op) and out[0] == out[1] == 0.0  # double root 0
    assert not lib.orc_solve_square_poly(1.0, 0.0, 1.0, op)


def _patch(q, u, v):
    q = np.asarray(q, np.float64)
    return (1 - u) * (1 - v) * q[0] + u * (1 - v) * q[1] + u * v * q[2] + (1 - u) * v * q[3]


def test_ray_quad_hits_the_aimed_patch_point(oracle):
    """Rays from outside aimed at P(u, v) of mildly twisted patches hit it:
    the returned (u, v) and t reproduce the point."""
    rng = np.random.default_rng(2)
    v, _, quads, _ = S.random_mixed_soup(0, 200, 0, seed=3, mu=0.0, twist=0.2)
    checked = 0
    for q in quads:
        Q = v[q].astype(np.float64)
        n = np.cross(Q[1] - Q[0], Q[3] - Q[0])
        n /= np.linalg.norm(n)
        u0, v0 = rng.uniform(0.05, 0.95, 2)
        target = _patch(Q, u0, v0)
        o = target + 3.0 * (n * rng.choice([-1, 1]) + 0.3 * rng.normal(size=3))
        d = target - o
        dist = np.linalg.norm(d)
        d /= dist
        hit, (t, u, vv) = ray_quad(oracle, o, d, Q)
        if hit and t < dist * (1 - 1e-3):
            continue  # the ray meets another part of the patch first (possible on twisted patches)
        assert hit
        assert abs(t - dist) <= 1e-4 * dist
        assert abs(u - u0) < 2e-3 and abs(vv - v0) < 2e-3
        assert np.linalg.norm(_patch(Q, u, vv) - (o + t * d)) < 1e-4 * dist
        checked += 1
    assert checked > 180


def test_ray_quad_misses_beside_a_planar_quad(oracle):
    Q = np.array([[0, 0, 0], [1, 0, 0], [1.2, 1, 0], [0, 1, 0]], np.float64)  # planar trapezoid
    rng = np.random.default_rng(4)
    for _ in range(200):
        u, v = rng.uniform(-1.0, 2.0, 2)
        if -0.05 <= u <= 1.05 and -0.05 <= v <= 1.05:
            continue
        target = _patch(Q, u, v)
        o = target + np.array([0.2, -0.1, 2.0])
        d = (target - o) / np.linalg.norm(target - o)
        hit, _ = ray_quad(oracle, o, d, Q)
        assert not hit
    # inside the planar quad: one root, the plane distance
    target = _patch(Q, 0.3, 0.6)
    o = target + np.array([0.0, 0.0, 2.0])
    hit, (t, u, v) = ray_quad(oracle, o, np.array([0.0, 0.0, -1.0]), Q)
    assert hit and abs(t - 2.0) < 1e-6 and abs(u - 0.3) < 1e-5 and abs(v - 0.6) < 1e-5


def test_ray_quad_takes_the_nearer_of_two_hits(oracle):
    """The patch (u, v, uv) is crossed twice by lines in direction (1, -1, 0):
    the nearer root of the double-precision quadratic, from either side."""
    Q = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 1], [0, 1, 0]], np.float64)
    d = np.array([1.0, -1.0, 0.0]) / np.sqrt(2.0)
    for z0 in (0.05, 0.2, 0.24):
        # (s, 1 - s, z0) on the patch: s (1 - s) = z0
        roots = np.sort(np.roots([1.0, -1.0, z0]).real)
        assert len(roots) == 2 and roots[0] > 0 and roots[1] < 1
        for sign in (1, -1):
            start = np.array([0.0, 1.0, z0]) if sign == 1 else np.array([1.0, 0.0, z0])
            o = start - sign * 3.0 * d
            hit, (t, u, v) = ray_quad(oracle, o, sign * d, Q)
            s_near = roots[0] if sign == 1 else 1.0 - roots[1]
            t_exp = 3.0 + s_near * np.sqrt(2.0)
            assert hit and abs(t - t_exp) < 1e-5, (z0, sign, t, t_exp)
            assert abs(u * v - z0) < 1e-5


def test_ray_sphere(oracle):
    rng = np.random.default_rng(5)
    for _ in range(300):
        c = rng.normal(size=3)
        r = rng.uniform(0.1, 2.0)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        dist = rng.uniform(1.5, 20.0) * r
        o = c - d * dist
        hit, t = ray_sphere(oracle, o, d, np.concatenate([c, [r]]))
        assert hit and abs(t - (dist - r)) <= 2e-6 * dist
        # from inside: the far root
        o2 = c + rng.uniform(-0.5, 0.5, 3) * r / np.sqrt(3)
        hit, t = ray_sphere(oracle, o2, d, np.concatenate([c, [r]]))
        oc = o2 - c
        b = np.dot(oc, d)
        t_far = -b + np.sqrt(b * b - (np.dot(oc, oc) - r * r))
        assert hit and abs(t - t_far) <= 2e-5 * r
        # behind the origin, and a clear miss
        hit, _ = ray_sphere(oracle, c + d * dist, d, np.concatenate([c, [r]]))
        assert not hit
        side = np.cross(d, rng.normal(size=3))
        side /= np.linalg.norm(side)
        hit, _ = ray_sphere(oracle, o + side * 1.01 * r, d, np.concatenate([c, [r]]))
        assert not hit


def test_bilerp_and_perpendicular(oracle):
    lib = _lib(oracle)
    x, xp = _f([1.5, -2.0, 3.25, 7.0])
    for (s, t), e in (((0, 0), 1.5), ((1, 0), -2.0), ((0, 1), 3.25), ((1, 1), 7.0)):
        assert lib.orc_bilerp(xp, s, t) == np.float32(e)
    assert abs(lib.orc_bilerp(xp, 0.5, 0.5) - np.mean([1.5, -2.0, 3.25, 7.0])) < 1e-6
    rng = np.random.default_rng(6)
    for v in list(rng.normal(size=(200, 3))) + [np.array([0, 0, 1.0]), np.array([1.0, 0, 0]), np.array([0, 1.0, 0])]:
        a, ap = _f(v)
        out, op = _f(np.zeros(3))
        lib.orc_perpendicular(ap, op)
        assert float(np.dot(a.astype(np.float64), out.astype(np.float64))) == 0.0
        assert np.linalg.norm(out) >= np.linalg.norm(a) / np.sqrt(3) * 0.99


# ------------------------------------------------------------------ LBVH
def _prim_boxes(v, t, q, sp):
    lo, hi = [], []
    if len(t):
        a, b = LC.tri_boxes(v, t)
        lo.append(a), hi.append(b)
    if len(q):
        P = v[q.astype(np.int64)]
        lo.append(P.min(axis=1)), hi.append(P.max(axis=1))
    if len(sp):
        lo.append(sp[:, :3] - sp[:, 3:]), hi.append(sp[:, :3] + sp[:, 3:])
    return np.concatenate(lo), np.concatenate(hi)


@pytest.mark.parametrize("nt,nq,ns,seed", [(300, 200, 100, 1), (0, 500, 0, 2), (0, 0, 700, 3), (1, 1, 1, 4),
                                           (0, 0, 1, 5), (2000, 1500, 1200, 6)])
def test_mixed_lbvh_invariants(oracle, nt, nq, ns, seed):
    v, t, q, sp = S.random_mixed_soup(nt, nq, ns, seed=seed)
    b = oracle.lbvh_build_mixed(v, t, q, sp)
    n = nt + nq + ns
    assert sorted(b["sorted_ids"].tolist()) == list(range(n))
    np.testing.assert_array_equal(b["bbox"], oracle.mixed_box(v, sp))
    if n == 1:
        assert b["num_internal"] == 0
        return
    B = b["boxes"].reshape(-1, 8)
    # leaf count codes by kind: triangle 1 << 20, quad 1 << 10, sphere 1
    ids = b["sorted_ids"]
    kind_code = np.where(ids < nt, 1 << 20, np.where(ids < nt + nq, 1 << 10, 1))
    np.testing.assert_array_equal(B[n - 1:, 3].view(np.int32), kind_code)
    # every internal node counts its primitives per kind (saturating at 1023)
    cnt = B[:n - 1, 3].view(np.int32)
    left, right = LC.children_from_parents(b["parents"], n)
    for x in range(n - 1):
        a, c = B[left[x], 3].view(np.int32), B[right[x], 3].view(np.int32)
        for sh in (20, 10, 0):
            assert (cnt[x] >> sh) & 0x3FF == min(((a >> sh) & 0x3FF) + ((c >> sh) & 0x3FF), 0x3FF)
    marks = np.diff(np.concatenate([[0], b["offsets"]]))
    np.testing.assert_array_equal(marks, LC.expected_marks(b["parents"], b["boxes"], n, LC.lod_capped,
                                                           LC.lod_leaf_cost))
    leaves = LC.walk_collapsed(b["nodes"], b["num_internal"], n)
    pos = 0
    for s, c in sorted((s, c) for s, c, _, _ in leaves):
        assert s == pos and 1 <= c
        pos += c
    assert pos == n
    plo, phi = _prim_boxes(v, t, q, sp)
    for s, c, lo, hi in leaves:
        p = ids[s:s + c]
        assert (lo <= plo[p].min(axis=0)).all() and (hi >= phi[p].max(axis=0)).all()


# ------------------------------------------------------------------ traversal
def mixed_geoms(seed):
    out = []
    for k, (nt, nq, ns) in enumerate(((400, 300, 200), (0, 60, 0), (0, 0, 80), (50, 0, 1), (0, 0, 1))):
        v, t, q, sp = S.random_mixed_soup(nt, nq, ns, seed=seed + k, mu=-2.5)
        out.append((v, t, q, sp))
    return out


@pytest.mark.parametrize("instances", [1, 12])
def test_mixed_trace_matches_bruteforce(oracle, instances):
    geoms = mixed_geoms(10)
    rng = np.random.default_rng(11)
    if instances == 1:
        inst_lod, w = np.array([0], np.int32), S.IDENTITY34.reshape(1, 12)
        rays = U.rays_toward(rng, 3000, (-1, 2), (0, 1))
    else:
        inst_lod = rng.integers(0, len(geoms), instances).astype(np.int32)
        w = U.random_transforms(instances, seed=12, spread=2.0)
        rays = U.rays_toward(rng, 3000)
    scene, _ = U.oracle_scene(oracle, geoms, inst_lod, w)
    t, inst, prim, uv = oracle.trace_closest(scene, rays)
    bt, binst, bprim = oracle.trace_brute(scene, rays)
    hit = inst >= 0
    assert hit.sum() > 300
    np.testing.assert_array_equal(inst, binst)
    np.testing.assert_array_equal(prim, bprim)
    if instances == 1:
        np.testing.assert_array_equal(t.view(np.uint32), bt.view(np.uint32))
    else:  # t passes object-space scalings in another order: ulp-level
        np.testing.assert_allclose(t[hit], bt[hit], rtol=1e-6, atol=0)
    occ = oracle.trace_any(scene, rays)
    np.testing.assert_array_equal(occ, hit.astype(np.int32))
    # every kind is hit
    kinds = set()
    for i, p in zip(inst[hit], prim[hit]):
        g = geoms[inst_lod[i]]
        kinds.add(0 if p < len(g[1]) else (1 if p < len(g[1]) + len(g[2]) else 2))
    assert kinds == {0, 1, 2}


# ------------------------------------------------------------------ lights
def _constant_emitter(le):
    return S.material(S.MAT_EMISSIVE, radiance=le, scale=1.0)


def test_quad_and_sphere_light_records_and_flux():
    """Records (AreaLightQuad: v0, v3 - v0, v1 - v0, v2 - v3 - (v1 - v0); the
    sphere's center and radius) and fluxes: a constant-radiance
    parallelogram's flux is 2 pi A sum(L), a sphere's 2 pi 4 pi r^2 sum(L)."""
    mats = np.concatenate([_constant_emitter((1.0, 2.0, 3.0)), _constant_emitter((0.5, 0.25, 4.0))])
    al = np.zeros(2, hip.AREA_LIGHT_EX_DTYPE)
    P = np.array([[0, 1, 0], [2, 1, 0], [3, 1, 1.5], [1, 1, 1.5]], np.float32)  # parallelogram, area 3
    al[0]["kind"], al[0]["points"], al[0]["material"], al[0]["instance"], al[0]["primitive"] = hip.AREA_QUAD, \
        P.reshape(-1), 0, 0, 5
    al[0]["uv"] = [0, 0, 1, 0, 1, 1, 0, 1]
    al[1]["kind"], al[1]["radius"], al[1]["material"], al[1]["instance"], al[1]["primitive"] = hip.AREA_SPHERE, 0.5, \
        1, 0, 9
    al[1]["points"][:3] = (4, 2, 1)
    info, mem, keys, codes = hip.light_tree_build(al, None, np.array([-1, -1, -1, 5, 5, 5], np.float32), mats)
    assert info["posLightCount"] == 2 and info["posRootType"] == 0xFFFF and info["mapCount"] == 2
    node = np.frombuffer(mem[:32].tobytes(), np.uint8)
    types = np.frombuffer(node[8:12].tobytes(), np.uint16)
    offs = [int(np.frombuffer(node[4:8].tobytes(), np.uint32)[0]), int(np.frombuffer(node[12:16].tobytes(),
                                                                                        np.uint32)[0])]
    fl = [float(np.frombuffer(node[0:4].tobytes(), np.float32)[0]), float(np.frombuffer(node[16:20].tobytes(),
                                                                                        np.float32)[0])]
    quad_side = list(types).index(3)
    sph_side = list(types).index(4)
    q = np.frombuffer(mem[offs[quad_side]:offs[quad_side] + 96].tobytes(), np.uint8)
    posV = np.frombuffer(q[:48].tobytes(), np.float32).reshape(4, 3)
    np.testing.assert_array_equal(posV, [P[0], P[3] - P[0], P[1] - P[0], (P[2] - P[3]) - (P[1] - P[0])])
    assert np.frombuffer(q[80:82].tobytes(), np.uint16)[0] == 0
    sr = np.frombuffer(mem[offs[sph_side]:offs[sph_side] + 32].tobytes(), np.uint8)
    np.testing.assert_array_equal(np.frombuffer(sr[:16].tobytes(), np.float32), [4, 2, 1, 0.5])
    assert np.frombuffer(sr[16:18].tobytes(), np.uint16)[0] == 1
    np.testing.assert_allclose(fl[quad_side], 2 * np.pi * 3.0 * 6.0, rtol=1e-5)
    np.testing.assert_allclose(fl[sph_side], 2 * np.pi * 4 * np.pi * 0.25 * 4.75, rtol=1e-5)
    assert sorted(keys.tolist()) == [5, 9]


def _pt_scene_mixed(O, lod, camera, point_lights=None, light_lod_box=None):
    """Oracle PtScene of one identity instance of a mixed LoD dict."""
    geoms = [(lod["verts"], lod["tris"], lod.get("quads"), lod.get("spheres"))]
    scene, lods = U.oracle_scene(O, geoms, np.array([0], np.int32), S.IDENTITY34.reshape(1, 12))
    box = lods[0]["box"]
    al = S.emissive_area_lights_ex(lod, lod["materials"])
    tree = hip.light_tree_build(al, point_lights, box, lod["materials"])
    bg = hip.background_record(hip.BACKGROUND_COLORED)
    flux, _ = hip.background_build(bg, None, box)
    info = tree[0].copy()
    info["backgroundFlux"] = np.float32(np.float32(flux[0] + flux[1]) + flux[2])
    return O.PtScene(scene, lod["normals"], lod["uvs"], lod["mat_idx"], lod["materials"], (info,) + tuple(tree[1:]),
                     box, camera, background=bg[0], background_flux=flux)


def _plane_with(emitters):
    """Lambert plane [-3, 3]^2 at y = 0 (two triangles, material 0) plus
    emitters: ("sphere", center, radius, L) or ("quad", 4 corners, L)."""
    mats = [S.material(S.MAT_LAMBERT, albedo=RHO)]
    v = [(-3, 0, 3), (3, 0, 3), (3, 0, -3), (-3, 0, -3)]
    n = [(0, 1, 0)] * 4
    uv = [(0, 0), (1, 0), (1, 1), (0, 1)]
    tris = [(0, 1, 2), (0, 2, 3)]
    quads, spheres, qm, sm = [], [], [], []
    for e in emitters:
        mats.append(_constant_emitter(e[-1]))
        if e[0] == "sphere":
            spheres.append(list(e[1]) + [e[2]])
            sm.append(len(mats) - 1)
        else:
            b = len(v)
            v += [tuple(p) for p in e[1]]
            n += [(0, -1, 0)] * 4
            uv += [(0, 0), (1, 0), (1, 1), (0, 1)]
            quads.append((b, b + 1, b + 2, b + 3))
            qm.append(len(mats) - 1)
    return dict(verts=np.array(v, np.float32), normals=np.array(n, np.float32), uvs=np.array(uv, np.float32),
                tris=np.array(tris, np.uint32), quads=np.array(quads, np.uint32).reshape(-1, 4),
                spheres=np.array(spheres, np.float32).reshape(-1, 4),
                mat_idx=np.array([0, 0] + qm + sm, np.uint16), materials=np.concatenate(mats))


def _stat_check(O, ps, expected, iterations, nee, max_len, seed=1, rel=2e-5):
    means, var, _ = O.pt_render_targets(ps, seed, 0, max_len, nee, iterations, [0, 1], variance=[0])
    hits = means[1][..., 1]
    m = means[0].astype(np.float64)
    v = var[0].astype(np.float64)
    n = m.shape[0] * m.shape[1]
    agg, exp = m.sum((0, 1)) / n, expected.sum((0, 1)) / n
    se = np.sqrt(v.sum((0, 1)) / iterations) / n
    dev = np.abs(agg - exp)
    return bool((dev <= 5.0 * se + rel * np.abs(exp)).all()), hits, \
        "mean %s expected %s se %s (%.1f se)" % (agg, exp, se, float((dev / se).max()))


SPHERE = ("sphere", (0.3, 0.75, -0.35), 0.12, (6.0, 4.0, 2.0))
TRAPEZOID = ("quad", [(-0.7, 0.6, -0.5), (-0.1, 0.6, -0.5), (0.1, 0.6, 0.0), (-0.6, 0.6, 0.0)], (3.0, 2.0, 1.0))


def _sphere_irradiance(x, center, radius):
    """A uniformly emitting sphere above the plane's horizon is a point source
    of intensity pi r^2 L: E = pi r^2 cos / d^2 per unit radiance."""
    d = np.asarray(center, np.float64) - x
    r2 = (d * d).sum(-1)
    return np.pi * radius * radius * (d[..., 1] / np.sqrt(r2)) / r2


@pytest.mark.parametrize("nee", [1, 2])
def test_sphere_light_irradiance(oracle, nee):
    cam = _camera(16)
    lod = _plane_with([SPHERE])
    ps = _pt_scene_mixed(oracle, lod, cam)
    x, wq = pixel_points(cam, *GAUSS3)
    E = _sphere_irradiance(x, SPHERE[1], SPHERE[2])[..., None] * np.asarray(SPHERE[3])
    expected = ((np.asarray(RHO) / np.pi * E) * wq[None, None, :, None]).sum(2)
    ok, hits, rep = _stat_check(oracle, ps, expected, 256, nee, 3)
    assert (hits <= 1).all(), "the camera must see the plane (primitives 0 and 1) only"
    assert ok, rep


@pytest.mark.parametrize("nee", [1, 2])
def test_trapezoid_quad_light_irradiance(oracle, nee):
    """A planar non-parallelogram quad light (its area density varies over the
    patch): Lambert's polygon formula."""
    cam = _camera(16)
    lod = _plane_with([TRAPEZOID])
    ps = _pt_scene_mixed(oracle, lod, cam)
    expected = expected_plane_radiance(cam, [(np.asarray(TRAPEZOID[1], np.float64), np.asarray(TRAPEZOID[2]))])
    ok, hits, rep = _stat_check(oracle, ps, expected, 256, nee, 3)
    assert (hits <= 1).all()
    assert ok, rep


def test_sphere_and_quad_lights_together(oracle):
    cam = _camera(16)
    lod = _plane_with([SPHERE, TRAPEZOID])
    ps = _pt_scene_mixed(oracle, lod, cam)
    x, wq = pixel_points(cam, *GAUSS3)
    E = _sphere_irradiance(x, SPHERE[1], SPHERE[2])[..., None] * np.asarray(SPHERE[3])
    expected = ((np.asarray(RHO) / np.pi * E) * wq[None, None, :, None]).sum(2)
    expected = expected + expected_plane_radiance(cam, [(np.asarray(TRAPEZOID[1], np.float64),
                                                         np.asarray(TRAPEZOID[2]))])
    ok, _, rep = _stat_check(oracle, ps, expected, 256, 2, 3)
    assert ok, rep


def test_sphere_connect_samples_the_visible_cap(oracle):
    """The sphere's next-event samples lie on the sphere, on the cap visible
    from the shading point, with creation pdf 1 / (2 pi (1 - r / d) r^2) (one
    light: choice probability 1), and light_pdf gives the same."""
    lod = _plane_with([SPHERE])
    ps = _pt_scene_mixed(oracle, lod, _camera(8))
    c, r = np.asarray(SPHERE[1], np.float64), SPHERE[2]
    rng = np.random.default_rng(7)
    for _ in range(200):
        pos = np.array([rng.uniform(-2, 2), 0.0, rng.uniform(-2, 2)], np.float32)
        u0, u1 = rng.random(2)
        p, pdf, _, leaf = oracle.light_connect(ps, 0, 1, 0, pos, float(u0), float(u1))
        p = np.asarray(p, np.float64)
        assert abs(np.linalg.norm(p - c) - r) < 1e-5
        assert np.dot(p - c, pos - p) >= -1e-6
        d = np.linalg.norm(pos - c)
        np.testing.assert_allclose(pdf, 1.0 / (2 * np.pi * (1 - r / d) * r * r), rtol=1e-5)
        lp = oracle.light_pdf(ps, 0, 2, pos)
        np.testing.assert_allclose(lp, pdf, rtol=1e-5)


def test_quad_light_pdf_is_the_local_area_density(oracle):
    """light_pdf of a hit on a quad light at patch coordinates (u, v) is the
    inverse of the patch's area density |dP/du x dP/dv| there (one light)."""
    lod = _plane_with([TRAPEZOID])
    ps = _pt_scene_mixed(oracle, lod, _camera(8))
    Q = np.asarray(TRAPEZOID[1], np.float64)
    for u, v in ((0.1, 0.2), (0.5, 0.5), (0.9, 0.3), (0.0, 1.0)):
        dPdu = (1 - v) * (Q[1] - Q[0]) + v * (Q[2] - Q[3])
        dPdv = (1 - u) * (Q[3] - Q[0]) + u * (Q[2] - Q[1])
        dens = np.linalg.norm(np.cross(dPdu, dPdv))
        lp = oracle.light_pdf(ps, 0, 2, np.zeros(3, np.float32), surface_params=(u, v))
        np.testing.assert_allclose(lp, 1.0 / dens, rtol=1e-5)
