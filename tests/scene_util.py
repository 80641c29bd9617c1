"""Scene helpers shared by the traversal tests (CPU oracle and GPU)."""
import numpy as np

from mufflon_amd import scenes as S


def random_transforms(n, seed, scale=(0.5, 2.0), spread=4.0):
    """worldToInstance matrices (n,12): random rotation * anisotropic scale + translation."""
    rng = np.random.default_rng(seed)
    w = np.zeros((n, 12), np.float32)
    for i in range(n):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        m = q * rng.uniform(scale[0], scale[1], 3)
        w[i, [0, 1, 2, 4, 5, 6, 8, 9, 10]] = m.reshape(-1)
        w[i, [3, 7, 11]] = rng.uniform(-spread, spread, 3)
    return w


def oracle_lod(O, verts, tris, parallel=False, quads=None, spheres=None):
    """Oracle LoD record; `parallel` builds with the OpenMP stages (same
    output). With quads (n x 4) or spheres (n x 4) the mixed build."""
    if (quads is not None and len(quads)) or (spheres is not None and len(spheres)):
        q = np.zeros((0, 4), np.uint32) if quads is None else quads
        sp = np.zeros((0, 4), np.float32) if spheres is None else spheres
        b = O.lbvh_build_mixed(verts, tris, q, sp)
        return dict(verts=verts, tris=tris, quads=q, spheres=sp, nodes=b["nodes"], num_internal=b["num_internal"],
                    prim_ids=b["sorted_ids"], box=b["bbox"]), b
    b = (O.lbvh_build_tris_omp if parallel else O.lbvh_build_tris)(verts, tris)
    return dict(verts=verts, tris=tris, nodes=b["nodes"], num_internal=b["num_internal"],
                prim_ids=b["sorted_ids"], box=b["bbox"]), b


def oracle_scene(O, lod_geoms, inst_lod, w, parallel=False):
    """Build LoD BVHs and (for >1 instance) the instance BVH with the oracle.
    lod_geoms: (verts, tris) or (verts, tris, quads, spheres) per LoD."""
    lods = [oracle_lod(O, g[0], g[1], parallel, *g[2:]) for g in lod_geoms]
    lods = [L for L, _ in lods]
    tlas = None
    if len(inst_lod) > 1:
        boxes = np.stack([L["box"] for L in lods])
        _, sb = O.instance_boxes(w, boxes, inst_lod)
        tb = O.lbvh_build_instances(w, boxes, inst_lod, sb)
        tlas = dict(nodes=tb["nodes"], num_internal=tb["num_internal"], prim_ids=tb["sorted_ids"])
    return O.Scene(lods, inst_lod, w, tlas), lods


def rays_toward(rng, n, origin_box=(-6, 6), target_box=(-3, 3), tmax=S.MAX_SCENE_SIZE):
    o = rng.uniform(origin_box[0], origin_box[1], (n, 3))
    tg = rng.uniform(target_box[0], target_box[1], (n, 3))
    d = tg - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmax
    r[:, 4:7] = d
    return r


def small_lods(seed=0):
    rng = np.random.default_rng(seed)
    v0, t0 = S.random_triangle_soup(3000, seed=seed + 1, mu=-3.0)
    v1, t1 = S.random_triangle_soup(500, seed=seed + 2, mu=-2.5)
    v1 = (v1 - 0.5) * 2.0
    v2, t2 = S.random_triangle_soup(1, seed=seed + 3, mu=-1.0)  # single-triangle LoD
    return [(v0.astype(np.float32), t0), (v1.astype(np.float32), t1), (v2.astype(np.float32), t2)], rng
