"""Closed-form test scenes shared by the estimator and adjoint tests: a
Lambert (or other) plane at y = 0 seen by a pinhole camera from above at an
angle (every pixel sees the plane), horizontal quad emitters, and the
per-pixel footprint quadrature of the camera's jittered directions."""
import numpy as np

import scene_util as U
from mufflon_amd import hip
from mufflon_amd import scenes as S

RHO = (0.55, 0.35, 0.2)
GAUSS3 = (np.array([0.5 - np.sqrt(0.15), 0.5, 0.5 + np.sqrt(0.15)]), np.array([5.0, 8.0, 5.0]) / 18.0)

def camera(size=24):
    pos = np.array([0.0, 0.45, 1.4])
    view = np.array([0.0, 0.0, 0.2]) - pos
    view /= np.linalg.norm(view)
    right = np.cross(view, [0.0, 1.0, 0.0])
    up = np.cross(right, view)
    up /= np.linalg.norm(up)
    return dict(position=pos.astype(np.float32), view_dir=view.astype(np.float32), up=up.astype(np.float32),
                tan_vfov=np.float32(np.tan(np.radians(14.0))), width=size, height=size)


def emitter_quad(mb, x0, x1, z0, z1, y, mat):
    """Horizontal quad emitter at height y facing down (-y)."""
    mb.add_quad((x0, y, z0), (x1, y, z0), (x1, y, z1), (x0, y, z1), (0, -1, 0), mat)
    return np.array([[x0, y, z0], [x1, y, z0], [x1, y, z1], [x0, y, z1]], np.float64)


def plane_scene(emitters=(), rho=RHO, plane_material=None):
    """Plane [-3,3]^2 at y = 0 (material 0: Lambert of albedo rho unless
    plane_material is given) and horizontal emitters given as (x0, x1, z0, z1,
    y, radiance); returns the mesh dict and the emitter polygons with their
    radiance."""
    mats = [S.material(S.MAT_LAMBERT, albedo=rho) if plane_material is None else plane_material]
    mb = S.MeshBuilder()
    mb.add_quad((-3, 0, 3), (3, 0, 3), (3, 0, -3), (-3, 0, -3), (0, 1, 0), 0)
    polys = []
    for k, (x0, x1, z0, z1, y, le) in enumerate(emitters):
        mats.append(S.material(S.MAT_EMISSIVE, radiance=le, scale=1.0))
        polys.append((emitter_quad(mb, x0, x1, z0, z1, y, k + 1), np.asarray(le, np.float64)))
    v, n, uv, t, m = mb.arrays()
    return dict(verts=v, normals=n, uvs=uv, tris=t, mat_idx=m, materials=np.concatenate(mats)), polys


def oracle_scene(O, c, camera, point_lights=None, background=None, pos_guide=False, textures=None):
    v, t = c["verts"], c["tris"]
    scene, _ = U.oracle_scene(O, [(v, t)], np.array([0], np.int32), S.IDENTITY34.reshape(1, 12))
    box = np.concatenate([v.min(0), v.max(0)]).astype(np.float32)
    al = S.emissive_area_lights(v, c["uvs"], t, c["mat_idx"], c["materials"])
    tree = hip.light_tree_build(al, point_lights, box, c["materials"])
    bg = background if background is not None else hip.background_record(hip.BACKGROUND_COLORED)
    flux, _ = hip.background_build(bg, None, box)
    info = tree[0].copy()
    info["backgroundFlux"] = np.float32(np.float32(flux[0] + flux[1]) + flux[2])
    tx = O.Textures(textures) if textures else None
    return O.PtScene(scene, c["normals"], c["uvs"], c["mat_idx"], c["materials"], (info,) + tuple(tree[1:]), box,
                     camera, textures=tx, background=bg[0], background_flux=flux, pos_guide=pos_guide)


def pixel_points(camera, nodes, weights):
    """Plane hit points (h, w, q, 3) of the jittered pixel directions at the
    quadrature nodes, and the weights (q,), in double."""
    W, H = camera["width"], camera["height"]
    view = camera["view_dir"].astype(np.float64)
    up = camera["up"].astype(np.float64)
    xa = np.cross(view, up)
    tv = float(camera["tan_vfov"])
    u0, u1 = np.meshgrid(nodes, nodes, indexing="ij")
    wq = np.outer(weights, weights).reshape(-1)
    px = np.arange(W)[None, :, None] + u0.reshape(-1)[None, None, :]
    py = np.arange(H)[:, None, None] + u1.reshape(-1)[None, None, :]
    cx = (px / W * 2.0 - 1.0) * tv * (W / H)
    cy = (py / H * 2.0 - 1.0) * tv
    cx, cy = np.broadcast_arrays(cx, cy)
    d = cx[..., None] * xa + cy[..., None] * up + view
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    o = camera["position"].astype(np.float64)
    t = -o[1] / d[..., 1]
    assert (t > 0).all(), "every pixel must see the plane"
    return o + d * t[..., None], wq


def polygon_irradiance(x, poly):
    """Irradiance per unit radiance at points x (..., 3) on the plane y = 0
    (normal +y) from a planar polygon seen entirely from its front side:
    Lambert's formula 1/2 sum_i theta_i (n . Gamma_i)."""
    tot = np.zeros(x.shape[:-1])
    for i in range(len(poly)):
        a = poly[i] - x
        b = poly[(i + 1) % len(poly)] - x
        la, lb = np.linalg.norm(a, axis=-1), np.linalg.norm(b, axis=-1)
        th = np.arccos(np.clip((a * b).sum(-1) / (la * lb), -1.0, 1.0))
        g = np.cross(a, b)
        tot += th * g[..., 1] / np.linalg.norm(g, axis=-1)
    return 0.5 * np.abs(tot)


def expected_plane_radiance(camera, polys, point_lights=(), rho=RHO):
    """rho/pi * E over each pixel's footprint (h, w, 3), double."""
    x, wq = pixel_points(camera, *GAUSS3)
    E = np.zeros(x.shape[:-1] + (3,))
    for poly, le in polys:
        E += polygon_irradiance(x, poly)[..., None] * le
    for pl in point_lights:
        d = pl["position"].astype(np.float64) - x
        r2 = (d * d).sum(-1)
        E += (d[..., 1] / np.sqrt(r2) / r2)[..., None] * pl["intensity"].astype(np.float64)
    L = np.asarray(rho) / np.pi * E
    return (L * wq[None, None, :, None]).sum(2)
