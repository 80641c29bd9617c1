"""Path-tracer scene setup shared by the oracle and GPU tests."""
import numpy as np

import scene_util as U
from mufflon_amd import hip
from mufflon_amd import scenes as S


def cornell(rings=16, segments=32, size=64):
    return S.cornell_box(rings, segments, size, size)


def scene_box(c):
    v = c["verts"]
    return np.concatenate([v.min(0), v.max(0)]).astype(np.float32)


def oracle_pt_scene(O, c, light_tree=None):
    v, t = c["verts"], c["tris"]
    scene, _ = U.oracle_scene(O, [(v, t)], np.array([0], np.int32), S.IDENTITY34.reshape(1, 12))
    box = scene_box(c)
    if light_tree is None:
        al = S.emissive_area_lights(v, c["uvs"], t, c["mat_idx"], c["materials"])
        light_tree = hip.light_tree_build(al, None, box, c["materials"])
    return O.PtScene(scene, c["normals"], c["uvs"], c["mat_idx"], c["materials"], light_tree, box, c["camera"])


def gpu_pt(ctx, c, seed=1, max_path_length=8, nee_count=1, min_path_length=0):
    lod = ctx.lod_create(c["verts"], c["tris"])
    ctx.lod_set_shading(lod, c["normals"], c["uvs"], c["mat_idx"])
    ctx.lbvh_build(lod, scene_box(c))
    sc = ctx.scene_create(np.array([lod], np.int32), S.IDENTITY34.reshape(1, 12))
    pt = ctx.pt_create(sc, c["materials"], c["camera"],
                       dict(seed=seed, min_path_length=min_path_length, max_path_length=max_path_length,
                            nee_count=nee_count))
    return pt


def _m34(M):
    return np.asarray(M, np.float64)[:3, :4].reshape(-1).astype(np.float32)


def _affine(R=np.eye(3), s=(1, 1, 1), t=(0, 0, 0)):
    M = np.eye(4)
    M[:3, :3] = np.asarray(R) @ np.diag(s)
    M[:3, 3] = t
    return M


def _rot_y(deg):
    a = np.radians(deg)
    return np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])


def cornell_instanced(size=48):
    """The Cornell box as separate LoDs placed by instances: walls (identity),
    a unit sphere instanced three times (scaled / moved), a unit cube twice
    (rotated, anisotropic scale) and the light quad (rotated half-turn, scaled)."""
    base = S.cornell_box(8, 16, size, size)
    mats = base["materials"]
    lods = []

    def mesh(fn):
        mb = S.MeshBuilder()
        fn(mb)
        v, n, uv, t, m = mb.arrays()
        return dict(verts=v, normals=n, uvs=uv, tris=t, mat_idx=m)

    lods.append(mesh(lambda mb: (
        mb.add_quad((-1, 0, 1), (1, 0, 1), (1, 0, -1), (-1, 0, -1), (0, 1, 0), 1),
        mb.add_quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1), (0, -1, 0), 1),
        mb.add_quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1), (0, 0, 1), 1),
        mb.add_quad((-1, 0, 1), (-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (1, 0, 0), 2),
        mb.add_quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1), (-1, 0, 0), 3))))
    lods.append(mesh(lambda mb: mb.add_uv_sphere((0, 0, 0), 1.0, 8, 16, 1)))
    lods.append(mesh(lambda mb: mb.add_box((0, 0, 0), (1, 1, 1), 0.0, 1)))
    lods.append(mesh(lambda mb: mb.add_quad((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1), (0, -1, 0), 0)))
    inst_to_world = [
        (0, _affine()),
        (1, _affine(s=(0.3, 0.3, 0.3), t=(0.1, 0.3, 0.45))),
        (1, _affine(s=(0.2, 0.25, 0.2), t=(-0.6, 0.25, 0.5))),
        (1, _affine(s=(0.15, 0.15, 0.15), t=(0.6, 1.3, -0.2))),
        (2, _affine(_rot_y(20), s=(0.5, 1.1, 0.5), t=(-0.4, 0.0, -0.35))),
        (2, _affine(_rot_y(-15), s=(0.45, 0.45, 0.5), t=(0.45, 0.0, -0.4))),
        (3, _affine(_rot_y(180), s=(0.3, 1.0, 0.2), t=(0.0, 1.99, 0.0))),
    ]
    inst_lod = np.array([i for i, _ in inst_to_world], np.int32)
    w2i = np.stack([_m34(np.linalg.inv(M)) for _, M in inst_to_world])
    return dict(lods=lods, inst_lod=inst_lod, world_to_inst=w2i, materials=mats, camera=base["camera"])


def gpu_pt_instanced(ctx, c, seed=1, max_path_length=8, nee_count=1):
    ids = []
    for L in c["lods"]:
        lod = ctx.lod_create(L["verts"], L["tris"])
        ctx.lod_set_shading(lod, L["normals"], L["uvs"], L["mat_idx"])
        ctx.lbvh_build(lod)
        ids.append(lod)
    sc = ctx.scene_create(np.array(ids, np.int32)[c["inst_lod"]], c["world_to_inst"])
    pt = ctx.pt_create(sc, c["materials"], c["camera"],
                       dict(seed=seed, max_path_length=max_path_length, nee_count=nee_count))
    return pt, sc


def oracle_pt_scene_instanced(O, c, light_tree, scene_box):
    geoms = [(L["verts"], L["tris"]) for L in c["lods"]]
    scene, _ = U.oracle_scene(O, geoms, c["inst_lod"], c["world_to_inst"])
    normals = np.concatenate([L["normals"] for L in c["lods"]])
    uvs = np.concatenate([L["uvs"] for L in c["lods"]])
    mat_idx = np.concatenate([L["mat_idx"] for L in c["lods"]])
    return O.PtScene(scene, normals, uvs, mat_idx, c["materials"], light_tree, scene_box, c["camera"])


def background_of(c, kind):
    """'none' (black), 'colored' or 'envmap' background record for scene c."""
    if kind == "envmap":
        return hip.background_record(hip.BACKGROUND_ENVMAP, scale=(0.5, 0.45, 0.6), envmap_texture=c["envmap"])
    if kind == "colored":
        return hip.background_record(hip.BACKGROUND_COLORED, color=(0.3, 0.4, 0.6), scale=0.8)
    return hip.background_record(hip.BACKGROUND_COLORED)


def gpu_pt_materials(ctx, c, background="envmap", seed=1, max_path_length=8, nee_count=1, pos_guide=False):
    ids = np.array([ctx.texture_create(t, s) for t, s in c["textures"]], np.int32)
    mats = c["materials"].copy()
    for f in ("albedoTexture", "roughnessTexture"):
        used = mats[f] >= 0
        mats[f][used] = ids[mats[f][used]]
    lod = ctx.lod_create(c["verts"], c["tris"])
    ctx.lod_set_shading(lod, c["normals"], c["uvs"], c["mat_idx"])
    ctx.lbvh_build(lod, scene_box(c))
    sc = ctx.scene_create(np.array([lod], np.int32), S.IDENTITY34.reshape(1, 12))
    pt = ctx.pt_create(sc, mats, c["camera"], dict(seed=seed, max_path_length=max_path_length, nee_count=nee_count,
                                                   nee_position_guide=pos_guide),
                       point_lights=c["point_lights"])
    bg = background_of(c, background)
    dev_bg = bg.copy()
    if int(bg["type"][0]) == hip.BACKGROUND_ENVMAP:
        dev_bg["envmapTexture"] = ids[c["envmap"]]
    ctx.pt_set_background(pt, dev_bg)
    return pt, sc, bg


def oracle_pt_scene_materials(O, c, background, light_tree=None, box=None, pos_guide=False):
    """Oracle scene of cornell_materials; without light_tree the host builders
    make the light tree (area + point lights) and the background."""
    v, t = c["verts"], c["tris"]
    scene, _ = U.oracle_scene(O, [(v, t)], np.array([0], np.int32), S.IDENTITY34.reshape(1, 12))
    box = scene_box(c) if box is None else box
    if light_tree is None:
        al = S.emissive_area_lights(v, c["uvs"], t, c["mat_idx"], c["materials"])
        light_tree = hip.light_tree_build(al, c["point_lights"], box, c["materials"])
    bg = background if isinstance(background, np.ndarray) else background_of(c, background)
    env = c["textures"][c["envmap"]][0] if int(bg["type"][0]) == hip.BACKGROUND_ENVMAP else None
    flux, sat = hip.background_build(bg, env, box)
    info = light_tree[0].copy()
    info["backgroundFlux"] = np.float32(np.float32(flux[0] + flux[1]) + flux[2])
    lt = (info,) + tuple(light_tree[1:])
    tx = O.Textures(c["textures"])
    return O.PtScene(scene, c["normals"], c["uvs"], c["mat_idx"], c["materials"], lt, box, c["camera"],
                     textures=tx, background=bg[0], background_sat=sat, background_flux=flux, pos_guide=pos_guide)


def gpu_pt_textured(ctx, c, seed=1, max_path_length=8, nee_count=1):
    """Path tracer of a scene with albedo textures (no background, no point
    lights); returns (pt, scene id, device texture ids)."""
    ids = np.array([ctx.texture_create(t, s) for t, s in c["textures"]], np.int32)
    mats = c["materials"].copy()
    for f in ("albedoTexture", "roughnessTexture"):
        used = mats[f] >= 0
        mats[f][used] = ids[mats[f][used]]
    lod = ctx.lod_create(c["verts"], c["tris"])
    ctx.lod_set_shading(lod, c["normals"], c["uvs"], c["mat_idx"])
    ctx.lbvh_build(lod, scene_box(c))
    sc = ctx.scene_create(np.array([lod], np.int32), S.IDENTITY34.reshape(1, 12))
    pt = ctx.pt_create(sc, mats, c["camera"], dict(seed=seed, max_path_length=max_path_length, nee_count=nee_count))
    return pt, sc, ids


def oracle_pt_scene_textured(O, c, textures=None, light_tree=None):
    v, t = c["verts"], c["tris"]
    scene, _ = U.oracle_scene(O, [(v, t)], np.array([0], np.int32), S.IDENTITY34.reshape(1, 12))
    box = scene_box(c)
    if light_tree is None:
        al = S.emissive_area_lights(v, c["uvs"], t, c["mat_idx"], c["materials"])
        light_tree = hip.light_tree_build(al, None, box, c["materials"])
    tx = O.Textures(textures if textures is not None else c["textures"])
    return O.PtScene(scene, c["normals"], c["uvs"], c["mat_idx"], c["materials"], light_tree, box, c["camera"],
                     textures=tx)


def oracle_pt_scene_general(O, sc, light_tree=None, box=None, background_scale=1.0, parallel=False):
    """Oracle scene of a workload scene dict (several LoDs, instances, point
    lights, textures, an optional polar env map): the oracle's own LBVHs and
    instance level; without light_tree / box the host builders make them.
    `parallel` builds the LoD BVHs with the OpenMP stages (same output)."""
    geoms = [(L["verts"], L["tris"], L.get("quads"), L.get("spheres")) for L in sc["lods"]]
    scene, lods = U.oracle_scene(O, geoms, sc["inst_lod"], sc["world_to_inst"], parallel)
    if box is None:
        boxes = np.stack([L["box"] for L in lods])
        _, box = O.instance_boxes(sc["world_to_inst"], boxes, sc["inst_lod"])
    box = np.asarray(box, np.float32)
    if light_tree is None:
        light_tree = hip.light_tree_build(None, sc.get("point_lights"), box, sc["materials"])
    tex = sc.get("textures", [])
    if sc.get("envmap") is not None:
        bg = hip.background_record(hip.BACKGROUND_ENVMAP, scale=background_scale, envmap_texture=sc["envmap"])
        env = tex[sc["envmap"]][0]
    else:
        bg = hip.background_record(hip.BACKGROUND_COLORED)
        env = None
    flux, sat = hip.background_build(bg, env, box)
    info = light_tree[0].copy()
    info["backgroundFlux"] = np.float32(np.float32(flux[0] + flux[1]) + flux[2])
    lt = (info,) + tuple(light_tree[1:])
    normals = np.concatenate([L["normals"] for L in sc["lods"]])
    uvs = np.concatenate([L["uvs"] for L in sc["lods"]])
    mat_idx = np.concatenate([L["mat_idx"] for L in sc["lods"]])
    return O.PtScene(scene, normals, uvs, mat_idx, sc["materials"], lt, box, sc["camera"], textures=O.Textures(tex),
                     background=bg[0], background_sat=sat, background_flux=flux)
