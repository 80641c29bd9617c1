"""Host light-tree builder (Mufflon's LightTreeBuilder layout) and the oracle
path tracer's closed-form cases. No device needed."""
import numpy as np
import pytest

import pt_util as P
from mufflon_amd import hip
from mufflon_amd import scenes as S

NODE = np.dtype([("lflux", "<f4"), ("loff", "<u4"), ("ltype", "<u2"), ("rtype", "<u2"), ("roff", "<u4"),
                 ("rflux", "<f4"), ("center", "<f4", 3)])
TRI = np.dtype([("posV", "<f4", 9), ("material", "<u2"), ("pad", "<u2"), ("uvV", "<f4", 6)])
POINT = np.dtype([("position", "<f4", 3), ("medium", "<u2"), ("pad", "<u2"), ("intensity", "<f4", 3),
                  ("pad1", "<f4")])
INTERNAL = 0xFFFF


def _random_lights(n_area, n_point, seed):
    rng = np.random.default_rng(seed)
    mats = np.concatenate([S.material(S.MAT_EMISSIVE, radiance=(2.0, 3.0, 4.0), scale=1.5),
                           S.material(S.MAT_EMISSIVE, radiance=(1.0, 0.5, 0.25)),
                           S.material(S.MAT_LAMBERT, albedo=(0.5, 0.5, 0.5))])
    al = np.zeros(n_area, hip.AREA_LIGHT_DTYPE)
    for i in range(n_area):
        c = rng.uniform(0.1, 0.9, 3)
        al[i]["points"] = (c + rng.normal(scale=0.05, size=(3, 3))).reshape(-1)
        al[i]["uv"] = rng.random(6)
        al[i]["material"] = i % 2
        al[i]["instance"] = i % 3
        al[i]["primitive"] = 1000 + i
    pl = np.zeros(n_point, hip.POINT_LIGHT_DTYPE)
    pl["position"] = rng.uniform(0.1, 0.9, (n_point, 3))
    pl["intensity"] = rng.uniform(0.5, 3.0, (n_point, 3))
    return al, pl, mats


def _walk(mem, info):
    """Yield (offset, type) of every leaf and check internal flux sums."""
    leaves = []
    stack = [(0, int(info["posRootType"]))]
    while stack:
        off, typ = stack.pop()
        if typ == INTERNAL:
            nd = np.frombuffer(mem[off:off + 32].tobytes(), NODE)[0]
            for o, t, f in ((nd["loff"], nd["ltype"], nd["lflux"]), (nd["roff"], nd["rtype"], nd["rflux"])):
                stack.append((int(o), int(t)))
                assert f > 0
        else:
            leaves.append((off, typ))
    return leaves


@pytest.mark.parametrize("n_area,n_point", [(1, 0), (2, 0), (3, 0), (5, 2), (13, 0), (0, 9), (32, 32)])
def test_tree_structure(n_area, n_point):
    al, pl, mats = _random_lights(n_area, n_point, n_area * 7 + n_point)
    box = np.array([0, 0, 0, 1, 1, 1], np.float32)
    info, mem, keys, codes = hip.light_tree_build(al, pl, box, mats)
    n = n_area + n_point
    assert info["posLightCount"] == n and info["mapCount"] == n_area
    h = int(np.floor(np.log2(n)))
    assert info["posInternalCount"] == (0 if n == 1 else (2 ** h - 1) + (n - 2 ** h))
    leaves = _walk(mem, info)
    assert len(leaves) == n
    assert sorted(t for _, t in leaves) == sorted([2] * n_area + [0] * n_point)
    # every primitive's code leads to its own triangle record
    for key, code in zip(keys, codes):
        off, typ = 0, int(info["posRootType"])
        while typ == INTERNAL:
            nd = np.frombuffer(mem[off:off + 32].tobytes(), NODE)[0]
            if code & 0x80000000:
                off, typ = int(nd["roff"]), int(nd["rtype"])
            else:
                off, typ = int(nd["loff"]), int(nd["ltype"])
            code = (int(code) << 1) & 0xFFFFFFFF
        assert typ == 2
        rec = np.frombuffer(mem[off:off + 64].tobytes(), TRI)[0]
        src = al[al["primitive"] == (int(key) & 0xFFFFFFFF)][0]
        assert np.array_equal(rec["posV"][:3], src["points"][:3])
        assert np.array_equal(rec["posV"][3:6], (src["points"][3:6] - src["points"][:3]).astype(np.float32))
    # root flux: float sum of the two root children (each a float sum below)
    if n > 1:
        nd = np.frombuffer(mem[:32].tobytes(), NODE)[0]
        assert info["posRootFlux"] == np.float32(nd["lflux"] + nd["rflux"])
        total = 0.0
        for off, typ in leaves:
            if typ == 0:
                r = np.frombuffer(mem[off:off + 32].tobytes(), POINT)[0]
                total += float(np.sum(r["intensity"].astype(np.float64) * 4 * np.pi))
            else:
                r = np.frombuffer(mem[off:off + 64].tobytes(), TRI)[0]
                e1, e2 = r["posV"][3:6].astype(np.float64), r["posV"][6:9].astype(np.float64)
                area = np.linalg.norm(np.cross(e1, e2)) / 2
                m = mats[r["material"]]
                total += float(np.sum(m["radiance"].astype(np.float64) * m["scale"]) * area * 2 * np.pi)
        assert abs(float(info["posRootFlux"]) - total) <= 1e-5 * total


def test_oracle_pt_sees_the_light_exactly(oracle):
    """Pixels whose camera ray hits a light triangle in every iteration carry
    exactly the light's radiance (no MIS on the first segment, no further
    bounce from an emitter)."""
    c = P.cornell(8, 16, 96)
    ps = P.oracle_pt_scene(oracle, c)
    rad, hit, nans = oracle.pt_render(ps, 1, 0, 8, 1, 4)
    light_prims = np.nonzero(c["mat_idx"] == 0)[0]
    on_light = np.isin(hit[..., 1], light_prims) & (hit[..., 0] == 0)
    assert on_light.sum() > 10 and nans == 0
    assert np.array_equal(rad[on_light], np.tile(np.float32([17.0, 12.0, 4.0]), (int(on_light.sum()), 1)))


def test_oracle_pt_is_deterministic_and_seeded(oracle):
    c = P.cornell(8, 16, 32)
    ps = P.oracle_pt_scene(oracle, c)
    a, _, _ = oracle.pt_render(ps, 1, 0, 8, 1, 2)
    b, _, _ = oracle.pt_render(ps, 1, 0, 8, 1, 2)
    d, _, _ = oracle.pt_render(ps, 2, 0, 8, 1, 2)
    assert np.array_equal(a, b) and not np.array_equal(a, d)


def test_oracle_pt_instanced_renders(oracle):
    """The instanced Cornell variant renders light through transformed lights."""
    c = P.cornell_instanced(32)
    # light gathered in world space as the device path does: instance 6 holds the light quad
    L = c["lods"][3]
    Minv = np.eye(4)
    Minv[:3, :] = c["world_to_inst"][6].reshape(3, 4)
    M = np.linalg.inv(Minv)
    al = np.zeros(2, hip.AREA_LIGHT_DTYPE)
    for j in range(2):
        t = L["tris"][j]
        p = (M[:3, :3] @ L["verts"][t].T.astype(np.float64)).T + M[:3, 3]
        al[j]["points"] = p.astype(np.float32).reshape(-1)
        al[j]["uv"] = L["uvs"][t].reshape(-1)
        al[j]["material"] = 0
        al[j]["instance"] = 6
        al[j]["primitive"] = j
    box = np.array([-1, 0, -1, 1, 2, 1], np.float32)
    lt = hip.light_tree_build(al, None, box, c["materials"])
    ps = P.oracle_pt_scene_instanced(oracle, c, lt, box)
    rad, hit, nans = oracle.pt_render(ps, 1, 0, 6, 1, 2)
    assert nans == 0 and rad.mean() > 0.01 and (hit[..., 0] == 6).any()


@pytest.mark.parametrize("background", ["none", "colored", "envmap"])
def test_oracle_pt_material_zoo(oracle, background):
    """Every material kind, textures, point lights and each background render
    without NaNs, and a lit background adds energy."""
    c = S.cornell_materials(8, 16, 24, 24)
    ps = P.oracle_pt_scene_materials(oracle, c, background)
    rad, hit, nans = oracle.pt_render(ps, 1, 0, 8, 1, 2)
    assert nans == 0 and np.isfinite(rad).all() and rad.mean() > 0.01
    if background != "none":
        base, _, _ = oracle.pt_render(P.oracle_pt_scene_materials(oracle, c, "none"), 1, 0, 8, 1, 2)
        assert rad.sum() > base.sum()


def test_oracle_targets_closed_forms(oracle):
    """Single-bounce Lambert scene: the guide-weighted first-hit targets follow
    from the geometry (depth of a pixel seeing the back wall straight on, unit
    normals, albedo of the hit material) and the Hit ID target equals the
    Radiance run's Hit ID."""
    c = P.cornell(8, 16, 16)
    ps = P.oracle_pt_scene(oracle, c)
    ref, nans = oracle.pt_render_targets(ps, 1, 0, 1, 1, 1, list(range(8)))
    rad, hit, _ = oracle.pt_render(ps, 1, 0, 1, 1, 1)
    assert nans == 0 and np.array_equal(ref[0], rad) and np.array_equal(ref[1], hit)
    # max path length 1: one segment, guide weight g = 1 / (1 + 1 / pdf_cam)
    n = ref[4]
    g = np.linalg.norm(n, axis=2)
    hitmask = hit[..., 0] >= 0
    depth = ref[3][..., 0]
    pos = ref[2]
    # position / depth / normal carry the same weight g: position = g * (eye + depth/g * dir)
    eye = np.array(c["camera"]["position"], np.float64)
    g = np.where(hitmask, g, 1.0)
    dist = np.linalg.norm(pos / g[..., None] - eye, axis=2)
    np.testing.assert_allclose(dist[hitmask], depth[hitmask] / g[hitmask], rtol=1e-4)
    alb = ref[6] / g[..., None]
    mats = c["materials"]
    prim = hit[..., 1].astype(int)
    expect = np.where(mats["type"][c["mat_idx"][prim]][..., None] == S.MAT_LAMBERT,
                      mats["albedo"][c["mat_idx"][prim]], 0.0)
    np.testing.assert_allclose(alb[hitmask], expect[hitmask], rtol=1e-5, atol=1e-6)


def _light_scene(oracle, al, pl, mats, pos_guide):
    """Oracle PtScene over a dummy one-triangle scene carrying the tree of
    the given lights (only the light fields are read by the hooks)."""
    import scene_util as U
    box = np.array([0, 0, 0, 1, 1, 1], np.float32)
    tree = hip.light_tree_build(al, pl, box, mats)
    v = np.array([[5, 5, 5], [6, 5, 5], [5, 6, 5]], np.float32)
    t = np.array([[0, 1, 2]], np.uint32)
    scene, _ = U.oracle_scene(oracle, [(v, t)], np.array([0], np.int32), S.IDENTITY34.reshape(1, 12))
    cam = dict(position=(0.5, 0.5, -2.0), view_dir=(0, 0, 1), up=(0, 1, 0), tan_vfov=0.5, width=4, height=4)
    ps = oracle.PtScene(scene, np.zeros((3, 3), np.float32), np.zeros((3, 2), np.float32), np.zeros(1, np.uint16),
                        mats, tree, box, cam, pos_guide=pos_guide)
    return ps, tree


def _leaf_probabilities(tree, ref, pos_guide):
    """Choice probability of every positional leaf (by byte offset), from the
    node records: guide_flux = lf / (lf + rf), or with the position guide
    lf / |cl - ref|^2 over the same sum (light_tree_sampling.hpp:58-71),
    with cl the child's center (node center, point position, triangle
    p0 + (e1 + e2) / 3). Also returns a relative error bound of the float
    evaluation: each factor q (pLeft or 1 - pLeft, rounded in float) adds
    about 2^-23 / q."""
    info, mem, _, _ = tree

    def center(off, typ):
        if typ == INTERNAL:
            return np.frombuffer(mem[off:off + 32].tobytes(), NODE)[0]["center"].astype(np.float64)
        if typ == 0:
            return np.frombuffer(mem[off:off + 32].tobytes(), POINT)[0]["position"].astype(np.float64)
        pv = np.frombuffer(mem[off:off + 64].tobytes(), TRI)[0]["posV"].astype(np.float64)
        return pv[:3] + (pv[3:6] + pv[6:9]) / 3

    probs = {}
    stack = [(0, int(info["posRootType"]), 1.0, 0.0)]
    ref = np.asarray(ref, np.float64)
    while stack:
        off, typ, p, err = stack.pop()
        if typ != INTERNAL:
            probs[off] = (typ, p, err)
            continue
        nd = np.frombuffer(mem[off:off + 32].tobytes(), NODE)[0]
        lf, rf = float(nd["lflux"]), float(nd["rflux"])
        if pos_guide:
            lf /= np.sum((center(int(nd["loff"]), int(nd["ltype"])) - ref) ** 2)
            rf /= np.sum((center(int(nd["roff"]), int(nd["rtype"])) - ref) ** 2)
        pl = lf / (lf + rf)
        eps = 2.0 ** -23
        stack.append((int(nd["loff"]), int(nd["ltype"]), p * pl, err + 4 * eps + eps / pl))
        stack.append((int(nd["roff"]), int(nd["rtype"]), p * (1 - pl), err + 4 * eps + eps / (1 - pl)))
    return probs


@pytest.mark.parametrize("pos_guide", [False, True])
def test_oracle_light_choice_probabilities(oracle, pos_guide):
    """NEE light choice through the tree, with and without the "NEE position
    guide": (i) the probability connect reports for the chosen light equals
    the product of the guide probabilities along its path (closed form from
    the node records); (ii) a uniform sweep of the choice variable lands on
    every light with that probability; (iii) light_pdf of a hit emitter
    triangle times its area equals the same probability (the hit-MIS walk
    agrees with the sampling walk)."""
    al, pl, mats = _random_lights(6, 5, 99)
    ps, tree = _light_scene(oracle, al, pl, mats, pos_guide)
    info, mem, keys, _ = tree
    n_sweep = 1 << 13
    for ref in ([0.5, 0.5, 0.5], [0.05, 0.9, 0.2], [1.5, -0.3, 0.7]):
        probs = _leaf_probabilities(tree, ref, pos_guide)
        hits = {}
        for i in range(n_sweep):
            _, pdf, _, leaf = oracle.light_connect(ps, i, n_sweep, 12345, ref)
            assert leaf in probs
            typ, p, err = probs[leaf]
            if typ == 0:
                got = pdf / 4294967296.0
            else:
                pv = np.frombuffer(mem[leaf:leaf + 64].tobytes(), TRI)[0]["posV"].astype(np.float64)
                got = pdf * np.linalg.norm(np.cross(pv[3:6], pv[6:9])) / 2
            assert abs(got - p) <= p * (2 * err + 1e-6), (ref, leaf, got, p, err)
            hits[leaf] = hits.get(leaf, 0) + 1
        for leaf, (_, p, _) in probs.items():
            assert abs(hits.get(leaf, 0) / n_sweep - p) <= 2.0 / n_sweep + 1e-6, (ref, leaf)
        for a in al:
            inst, prim = int(a["instance"]), int(a["primitive"])
            lp = oracle.light_pdf(ps, inst, prim, ref)
            key = (inst << 32) | prim
            j = int(np.nonzero(keys == key)[0][0])
            # the leaf of this triangle: the record whose first point matches
            leaf = [o for o, (t, _, _) in probs.items()
                    if t == 2 and np.array_equal(np.frombuffer(mem[o:o + 12].tobytes(), "<f4"), a["points"][:3])]
            assert len(leaf) == 1, j
            pts = a["points"].reshape(3, 3).astype(np.float32)
            area = np.linalg.norm(np.cross((pts[1] - pts[0]).astype(np.float64), (pts[2] - pts[0]).astype(np.float64))) / 2
            _, p, err = probs[leaf[0]]
            assert abs(lp * area - p) <= p * (2 * err + 1e-6), (ref, j)
    if pos_guide:
        # the guide changes the choice with the reference position
        a = _leaf_probabilities(tree, [0.05, 0.9, 0.2], True)
        b = _leaf_probabilities(tree, [0.9, 0.1, 0.8], True)
        assert max(abs(a[k][1] - b[k][1]) for k in a) > 0.05
