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
