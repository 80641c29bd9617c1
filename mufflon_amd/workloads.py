"""Device set-up of the procedural workloads (scenes of mufflon_amd.scenes)
through the C-ABI: LoDs with shading attributes, LBVHs, the instance level,
textures, the path tracer and its background."""
import numpy as np

from mufflon_amd import hip


def upload_lods(ctx, lods, build=True):
    ids = []
    for L in lods:
        if len(L.get("quads", ())) or len(L.get("spheres", ())):
            lod = ctx.lod_create_mixed(L["verts"], L["tris"], L.get("quads"), L.get("spheres"))
        else:
            lod = ctx.lod_create(L["verts"], L["tris"])
        ctx.lod_set_shading(lod, L["normals"], L["uvs"], L["mat_idx"])
        if build:
            ctx.lbvh_build(lod)
        ids.append(lod)
    return ids


def path_tracer(ctx, scene, seed=1, max_path_length=16, nee_count=1, background_scale=1.0):
    """Path tracer over a scene dict with 'lods', 'inst_lod', 'world_to_inst',
    'materials', 'point_lights', 'textures', 'envmap', 'camera'. Returns
    (pt id, scene id, texture ids)."""
    tex = np.array([ctx.texture_create(t, s) for t, s in scene.get("textures", [])], np.int32)
    mats = scene["materials"].copy()
    for f in ("albedoTexture", "roughnessTexture"):
        used = mats[f] >= 0
        mats[f][used] = tex[mats[f][used]]
    lod_ids = np.array(upload_lods(ctx, scene["lods"]), np.int32)
    sc = ctx.scene_create(lod_ids[scene["inst_lod"]], scene["world_to_inst"])
    pt = ctx.pt_create(sc, mats, scene["camera"],
                       dict(seed=seed, max_path_length=max_path_length, nee_count=nee_count),
                       point_lights=scene.get("point_lights"))
    if scene.get("envmap") is not None:
        ctx.pt_set_background(pt, hip.background_record(hip.BACKGROUND_ENVMAP, scale=background_scale,
                                                        envmap_texture=int(tex[scene["envmap"]])))
    return pt, sc, tex


def uneven_batch(ctx, count=48, seed0=100, use_graph=True, scenes=None, with_scene_ids=False):
    """The uneven batch: `count` procedural scenes (scenes.batch_scene) as path
    tracers and one batch over them. Returns (batch id, path tracer ids,
    texture ids per scene, scene dicts), plus the scene ids with
    `with_scene_ids`."""
    from mufflon_amd import scenes as S
    if scenes is None:
        scenes = [S.batch_scene(i, seed0) for i in range(count)]
    pts, texs, sids = [], [], []
    for sc in scenes:
        pt, sid, tex = path_tracer(ctx, sc, max_path_length=sc["max_path_length"])
        pts.append(pt)
        texs.append(tex)
        sids.append(sid)
    batch = ctx.batch_create(pts, use_graph)
    if with_scene_ids:
        return batch, pts, texs, scenes, sids
    return batch, pts, texs, scenes
