"""Oracle — test infrastructure only.

ctypes access to the CPU restatement of Mufflon's CPU path
(oracle/build/liboracle.so, built by `make -C oracle`). Only the tests, the
smoke check and the CPU-baseline leg of the benchmark use it, as the checker
or as the timed CPU baseline; no product code path calls it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liboracle.so")
_lib = None

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class OrcScene(C.Structure):
    _fields_ = [
        ("numLods", C.c_int32),
        ("lodVertOffset", i32p), ("lodTriOffset", i32p), ("lodNodeOffset", i32p), ("lodNumInternal", i32p),
        ("verts", f32p), ("tris", u32p), ("nodes", C.c_void_p), ("primIds", i32p), ("lodBoxes", f32p),
        ("numInstances", C.c_int32), ("worldToInst", f32p), ("instLod", i32p),
        ("tlasNodes", C.c_void_p), ("tlasNumInternal", C.c_int32), ("tlasPrimIds", i32p),
        ("lodQuadOffset", i32p), ("quads", u32p), ("lodSphereOffset", i32p), ("spheres", f32p),
        ("lodPrimOffset", i32p),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        lib.orc_morton_point.argtypes = [f32p, f32p]
        lib.orc_morton_point.restype = C.c_uint64
        lib.orc_part_by_two21.argtypes = [C.c_uint32]
        lib.orc_part_by_two21.restype = C.c_uint64
        lib.orc_part_by_two21_batch.argtypes = [u32p, C.c_int64, u64p]
        lib.orc_part_by_two21_batch.restype = None
        lib.orc_morton_points.argtypes = [f32p, C.c_int64, f32p, u64p]
        lib.orc_morton_points.restype = None
        lib.orc_lbvh_build_tris.argtypes = [f32p, u32p, C.c_int32, f32p, u64p, u64p, i32p, i32p, f32p, i32p,
                                            C.c_void_p, i32p]
        lib.orc_lbvh_build_mixed.argtypes = [f32p, u32p, C.c_int32, u32p, C.c_int32, f32p, C.c_int32, f32p, u64p,
                                             i32p, i32p, f32p, i32p, C.c_void_p, i32p]
        lib.orc_lbvh_build_instances.argtypes = [f32p, f32p, i32p, C.c_int32, f32p, u64p, i32p, i32p, f32p, i32p,
                                                 C.c_void_p, i32p]
        lib.orc_instance_boxes.argtypes = [f32p, f32p, i32p, C.c_int32, f32p, f32p]
        lib.orc_karras.argtypes = [u64p, C.c_int32, i32p]
        lib.orc_trace_closest.argtypes = [C.POINTER(OrcScene), f32p, C.c_int32, f32p, i32p, i32p, f32p, u32p, u32p]
        lib.orc_trace_any.argtypes = [C.POINTER(OrcScene), f32p, C.c_int32, i32p]
        lib.orc_trace_brute.argtypes = [C.POINTER(OrcScene), f32p, C.c_int32, f32p, i32p, i32p]
        for f in ("orc_lbvh_build_tris", "orc_lbvh_build_instances", "orc_instance_boxes", "orc_karras",
                  "orc_trace_closest", "orc_trace_any", "orc_trace_brute"):
            getattr(lib, f).restype = C.c_int
        _lib = lib
    return _lib


def _p(a, t):
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("oracle arrays must be C-contiguous")
    return a.ctypes.data_as(t)


def part_by_two21(x):
    x = np.ascontiguousarray(x, np.uint32)
    out = np.zeros(len(x), np.uint64)
    load().orc_part_by_two21_batch(_p(x, u32p), len(x), _p(out, u64p))
    return out


def morton_points(points, bbox):
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    bb = np.ascontiguousarray(bbox, np.float32)
    out = np.zeros(len(p), np.uint64)
    load().orc_morton_points(_p(p, f32p), len(p), _p(bb, f32p), _p(out, u64p))
    return out


def vertex_box(verts):
    v = np.asarray(verts, np.float32).reshape(-1, 3)
    return np.concatenate([v.min(axis=0), v.max(axis=0)]).astype(np.float32)


def lbvh_build_tris(verts, tris, bbox=None):
    """Serial LBVH build; returns a dict of every stage."""
    lib = load()
    v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
    n = len(t)
    bb = vertex_box(v) if bbox is None else np.ascontiguousarray(bbox, np.float32)
    uk = np.zeros(n, np.uint64)
    sk = np.zeros(n, np.uint64)
    ids = np.zeros(n, np.int32)
    par = np.zeros(max(2 * n - 1, 1), np.int32)
    boxes = np.zeros(max(2 * n - 1, 1) * 8, np.float32)
    offs = np.zeros(max(n - 1, 1), np.int32)
    nodes = np.zeros(max(n - 1, 1) * 64, np.uint8)
    nint = C.c_int32()
    rc = lib.orc_lbvh_build_tris(_p(v, f32p), _p(t, u32p), n, _p(bb, f32p), _p(uk, u64p), _p(sk, u64p),
                                 _p(ids, i32p), _p(par, i32p), _p(boxes, f32p), _p(offs, i32p),
                                 nodes.ctypes.data, C.byref(nint))
    assert rc == 0
    return dict(unsorted_keys=uk, sorted_keys=sk, sorted_ids=ids, parents=par, boxes=boxes, offsets=offs,
                nodes=nodes[: nint.value * 64].copy(), num_internal=nint.value, bbox=bb)


def mixed_box(verts, spheres):
    """LoD box of Lod::get_bounding_box: the vertices' box joined with every
    sphere's center -+ radius."""
    parts = []
    v = np.asarray(verts, np.float32).reshape(-1, 3)
    if len(v):
        parts += [v.min(axis=0), v.max(axis=0)]
    sp = np.asarray(spheres, np.float32).reshape(-1, 4)
    if len(sp):
        parts += [(sp[:, :3] - sp[:, 3:]).min(axis=0), (sp[:, :3] + sp[:, 3:]).max(axis=0)]
    lo = np.minimum.reduce(parts[0::2])
    hi = np.maximum.reduce(parts[1::2])
    return np.concatenate([lo, hi]).astype(np.float32)


def lbvh_build_mixed(verts, tris, quads, spheres, bbox=None):
    """Serial LBVH build of a LoD of triangles, quads (4 vertex indices) and
    spheres (center, radius); primitive ids run triangles, quads, spheres."""
    lib = load()
    v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
    q = np.ascontiguousarray(quads, np.uint32).reshape(-1, 4)
    sp = np.ascontiguousarray(spheres, np.float32).reshape(-1, 4)
    n = len(t) + len(q) + len(sp)
    bb = mixed_box(v, sp) if bbox is None else np.ascontiguousarray(bbox, np.float32)
    sk = np.zeros(n, np.uint64)
    ids = np.zeros(n, np.int32)
    par = np.zeros(max(2 * n - 1, 1), np.int32)
    boxes = np.zeros(max(2 * n - 1, 1) * 8, np.float32)
    offs = np.zeros(max(n - 1, 1), np.int32)
    nodes = np.zeros(max(n - 1, 1) * 64, np.uint8)
    nint = C.c_int32()
    keep = [np.zeros(4, np.uint32), np.zeros(4, np.float32), np.zeros(3, np.float32)]
    rc = lib.orc_lbvh_build_mixed(_p(v if len(v) else keep[2], f32p), _p(t if len(t) else keep[0], u32p), len(t),
                                  _p(q if len(q) else keep[0], u32p), len(q), _p(sp if len(sp) else keep[1], f32p),
                                  len(sp), _p(bb, f32p), _p(sk, u64p), _p(ids, i32p), _p(par, i32p),
                                  _p(boxes, f32p), _p(offs, i32p), nodes.ctypes.data, C.byref(nint))
    assert rc == 0
    return dict(sorted_keys=sk, sorted_ids=ids, parents=par, boxes=boxes, offsets=offs,
                nodes=nodes[: nint.value * 64].copy(), num_internal=nint.value, bbox=bb)


def lbvh_build_tris_omp(verts, tris, bbox=None):
    """The OpenMP-parallel LBVH build (the reference's parallel CPU path for
    LoDs of 1000+ primitives); returns the same dict as lbvh_build_tris
    without the unsorted keys."""
    lib = load()
    v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
    n = len(t)
    bb = vertex_box(v) if bbox is None else np.ascontiguousarray(bbox, np.float32)
    sk = np.zeros(n, np.uint64)
    ids = np.zeros(n, np.int32)
    par = np.zeros(max(2 * n - 1, 1), np.int32)
    boxes = np.zeros(max(2 * n - 1, 1) * 8, np.float32)
    offs = np.zeros(max(n - 1, 1), np.int32)
    nodes = np.zeros(max(n - 1, 1) * 64, np.uint8)
    nint = C.c_int32()
    lib.orc_lbvh_build_tris_omp.argtypes = [f32p, u32p, C.c_int32, f32p, u64p, i32p, i32p, f32p, i32p, C.c_void_p,
                                            C.POINTER(C.c_int32)]
    rc = lib.orc_lbvh_build_tris_omp(_p(v, f32p), _p(t, u32p), n, _p(bb, f32p), _p(sk, u64p), _p(ids, i32p),
                                     _p(par, i32p), _p(boxes, f32p), _p(offs, i32p), nodes.ctypes.data,
                                     C.byref(nint))
    assert rc == 0
    return dict(sorted_keys=sk, sorted_ids=ids, parents=par, boxes=boxes, offsets=offs,
                nodes=nodes[: nint.value * 64].copy(), num_internal=nint.value, bbox=bb)


def instance_boxes(world_to_inst, obj_boxes, inst_lod):
    lib = load()
    w = np.ascontiguousarray(world_to_inst, np.float32).reshape(-1, 12)
    ob = np.ascontiguousarray(obj_boxes, np.float32).reshape(-1, 6)
    il = np.ascontiguousarray(inst_lod, np.int32)
    out = np.zeros((len(il), 6), np.float32)
    sb = np.zeros(6, np.float32)
    lib.orc_instance_boxes(_p(w, f32p), _p(ob, f32p), _p(il, i32p), len(il), _p(out, f32p), _p(sb, f32p))
    return out, sb


def lbvh_build_instances(world_to_inst, obj_boxes, inst_lod, scene_box):
    lib = load()
    w = np.ascontiguousarray(world_to_inst, np.float32).reshape(-1, 12)
    ob = np.ascontiguousarray(obj_boxes, np.float32).reshape(-1, 6)
    il = np.ascontiguousarray(inst_lod, np.int32)
    sb = np.ascontiguousarray(scene_box, np.float32)
    n = len(il)
    sk = np.zeros(n, np.uint64)
    ids = np.zeros(n, np.int32)
    par = np.zeros(max(2 * n - 1, 1), np.int32)
    boxes = np.zeros(max(2 * n - 1, 1) * 8, np.float32)
    offs = np.zeros(max(n - 1, 1), np.int32)
    nodes = np.zeros(max(n - 1, 1) * 64, np.uint8)
    nint = C.c_int32()
    rc = lib.orc_lbvh_build_instances(_p(w, f32p), _p(ob, f32p), _p(il, i32p), n, _p(sb, f32p), _p(sk, u64p),
                                      _p(ids, i32p), _p(par, i32p), _p(boxes, f32p), _p(offs, i32p),
                                      nodes.ctypes.data, C.byref(nint))
    assert rc == 0
    return dict(sorted_keys=sk, sorted_ids=ids, parents=par, boxes=boxes, offsets=offs,
                nodes=nodes[: nint.value * 64].copy(), num_internal=nint.value)


def karras(sorted_keys):
    lib = load()
    k = np.ascontiguousarray(sorted_keys, np.uint64)
    par = np.zeros(2 * len(k) - 1, np.int32)
    assert lib.orc_karras(_p(k, u64p), len(k), _p(par, i32p)) == 0
    return par


class Scene:
    """Flat multi-LoD, multi-instance scene for the oracle tracer."""

    def __init__(self, lods, inst_lod, world_to_inst, tlas=None):
        # lods: list of dict(verts, tris, nodes, num_internal, prim_ids, box),
        # optionally quads (n x 4 vertex indices) and spheres (n x 4: center, radius)
        self._keep = []
        vo, to, no = [0], [0], [0]
        for L in lods:
            vo.append(vo[-1] + len(L["verts"]))
            to.append(to[-1] + len(L["tris"]))
            no.append(no[-1] + L["num_internal"])
        self.verts = np.ascontiguousarray(np.concatenate([np.asarray(L["verts"], np.float32).reshape(-1, 3)
                                                          for L in lods]))
        self.tris = np.ascontiguousarray(np.concatenate([np.asarray(L["tris"], np.uint32).reshape(-1, 3)
                                                         for L in lods]))
        self.nodes = np.ascontiguousarray(np.concatenate([L["nodes"] for L in lods] + [np.zeros(64, np.uint8)]))
        self.prim_ids = np.ascontiguousarray(np.concatenate([L["prim_ids"] for L in lods]).astype(np.int32))
        self.boxes = np.ascontiguousarray(np.concatenate([L["box"] for L in lods]).astype(np.float32))
        self.vo = np.array(vo, np.int32)
        self.to = np.array(to, np.int32)
        self.no = np.array(no, np.int32)
        self.ni = np.array([L["num_internal"] for L in lods], np.int32)
        self.w = np.ascontiguousarray(world_to_inst, np.float32).reshape(-1, 12)
        self.il = np.ascontiguousarray(inst_lod, np.int32)
        if tlas is None:
            tlas = dict(nodes=np.zeros(64, np.uint8), num_internal=0, prim_ids=np.zeros(1, np.int32))
        self.tn = np.ascontiguousarray(tlas["nodes"] if len(tlas["nodes"]) else np.zeros(64, np.uint8))
        self.tp = np.ascontiguousarray(tlas["prim_ids"], np.int32)
        s = OrcScene()
        s.numLods = len(lods)
        s.lodVertOffset = _p(self.vo, i32p); s.lodTriOffset = _p(self.to, i32p)
        s.lodNodeOffset = _p(self.no, i32p); s.lodNumInternal = _p(self.ni, i32p)
        s.verts = _p(self.verts, f32p); s.tris = _p(self.tris, u32p)
        s.nodes = self.nodes.ctypes.data; s.primIds = _p(self.prim_ids, i32p); s.lodBoxes = _p(self.boxes, f32p)
        s.numInstances = len(self.il); s.worldToInst = _p(self.w, f32p); s.instLod = _p(self.il, i32p)
        s.tlasNodes = self.tn.ctypes.data; s.tlasNumInternal = int(tlas["num_internal"])
        s.tlasPrimIds = _p(self.tp, i32p)
        if any(len(L.get("quads", ())) or len(L.get("spheres", ())) for L in lods):
            qo, so, po = [0], [0], [0]
            for L in lods:
                nq, ns = len(L.get("quads", ())), len(L.get("spheres", ()))
                qo.append(qo[-1] + nq)
                so.append(so[-1] + ns)
                po.append(po[-1] + len(L["tris"]) + nq + ns)
            self.qo, self.so, self.po = (np.array(x, np.int32) for x in (qo, so, po))
            self.quads = np.ascontiguousarray(np.concatenate(
                [np.asarray(L.get("quads", np.zeros((0, 4))), np.uint32).reshape(-1, 4) for L in lods]
                + [np.zeros((1, 4), np.uint32)]))
            self.spheres = np.ascontiguousarray(np.concatenate(
                [np.asarray(L.get("spheres", np.zeros((0, 4))), np.float32).reshape(-1, 4) for L in lods]
                + [np.zeros((1, 4), np.float32)]))
            s.lodQuadOffset = _p(self.qo, i32p); s.quads = _p(self.quads, u32p)
            s.lodSphereOffset = _p(self.so, i32p); s.spheres = _p(self.spheres, f32p)
            s.lodPrimOffset = _p(self.po, i32p)
        self.s = s


def trace_closest(scene, rays, counts=False):
    lib = load()
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = len(r)
    t = np.zeros(n, np.float32)
    inst = np.zeros(n, np.int32)
    prim = np.zeros(n, np.int32)
    bary = np.zeros((n, 2), np.float32)
    nc = np.zeros(n, np.uint32) if counts else None
    pc = np.zeros(n, np.uint32) if counts else None
    lib.orc_trace_closest(C.byref(scene.s), _p(r, f32p), n, _p(t, f32p), _p(inst, i32p), _p(prim, i32p),
                          _p(bary, f32p), _p(nc, u32p) if counts else None, _p(pc, u32p) if counts else None)
    if counts:
        return t, inst, prim, bary, nc, pc
    return t, inst, prim, bary


def trace_any(scene, rays):
    lib = load()
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    o = np.zeros(len(r), np.int32)
    lib.orc_trace_any(C.byref(scene.s), _p(r, f32p), len(r), _p(o, i32p))
    return o


def trace_brute(scene, rays):
    lib = load()
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = len(r)
    t = np.zeros(n, np.float32)
    inst = np.zeros(n, np.int32)
    prim = np.zeros(n, np.int32)
    lib.orc_trace_brute(C.byref(scene.s), _p(r, f32p), n, _p(t, f32p), _p(inst, i32p), _p(prim, i32p))
    return t, inst, prim


class OrcTexture(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("channels", C.c_int32), ("sampling", C.c_int32),
                ("texels", f32p)]


class OrcBackground(C.Structure):
    _fields_ = [("type", C.c_int32), ("color", C.c_float * 3), ("scale", C.c_float * 3), ("envmapTexture", C.c_int32),
                ("sat", f32p), ("flux", C.c_float * 3)]


class OrcPtScene(C.Structure):
    _fields_ = [
        ("geo", OrcScene),
        ("normals", f32p), ("uvs", f32p), ("matIdx", C.c_void_p), ("materials", C.c_void_p),
        ("numMaterials", C.c_int32),
        ("lightMem", C.c_void_p), ("posRootFlux", C.c_float), ("posRootType", C.c_uint32),
        ("posInternalCount", C.c_uint32), ("posLightCount", C.c_uint32), ("dirRootFlux", C.c_float),
        ("bgFlux", C.c_float), ("mapKeys", u64p), ("mapCodes", u32p), ("mapCount", C.c_int32),
        ("sceneBox", C.c_float * 6), ("camPos", C.c_float * 3), ("tanVFov", C.c_float),
        ("camDir", C.c_float * 3), ("camUp", C.c_float * 3), ("width", C.c_int32), ("height", C.c_int32),
        ("textures", C.POINTER(OrcTexture)), ("numTextures", C.c_int32), ("background", OrcBackground),
        ("posGuide", C.c_int32),
    ]


class Textures:
    """Texture table in the oracle layout: each entry (texels (h, w, c) float32,
    sampling 0 nearest / 1 linear) becomes RGBA float texels."""

    def __init__(self, textures=()):
        self.data = []
        self.arr = (OrcTexture * max(len(textures), 1))()
        for i, (texels, sampling) in enumerate(textures):
            t = np.asarray(texels, np.float32)
            if t.ndim == 2:
                t = t[:, :, None]
            h, w, c = t.shape
            rgba = np.zeros((h, w, 4), np.float32)
            rgba[:, :, :c] = t
            self.data.append(rgba)
            self.arr[i].width, self.arr[i].height, self.arr[i].channels = w, h, c
            self.arr[i].sampling = int(sampling)
            self.arr[i].texels = _p(rgba, f32p)
        self.count = len(textures)


def texture_sample(textures, index, u, v):
    lib = load()
    out = np.zeros(4, np.float32)
    lib.orc_texture_sample.argtypes = [C.POINTER(OrcTexture), C.c_float, C.c_float, f32p]
    lib.orc_texture_sample(C.byref(textures.arr[index]), u, v, _p(out, f32p))
    return out


def _bsdf_lib():
    lib = load()
    if not hasattr(lib, "_bsdf_ready"):
        lib.orc_bsdf_sample.argtypes = [C.c_void_p, C.POINTER(OrcTexture), C.c_float, C.c_float, f32p, C.c_float,
                                        C.c_float, C.c_uint64, f32p]
        lib.orc_bsdf_evaluate.argtypes = [C.c_void_p, C.POINTER(OrcTexture), C.c_float, C.c_float, f32p, f32p, f32p]
        lib.orc_bsdf_evaluate_batch.argtypes = [C.c_void_p, C.POINTER(OrcTexture), C.c_float, C.c_float, C.c_int32,
                                                f32p, f32p, f32p]
        lib.orc_eval_ndf.argtypes = [C.c_int32, C.c_float, C.c_float, f32p]
        lib.orc_eval_ndf.restype = C.c_float
        lib.orc_fresnel_dielectric.argtypes = [C.c_float, C.c_float, C.c_float, f32p]
        lib.orc_fresnel_dielectric.restype = C.c_float
        lib.orc_background_sample.argtypes = [C.POINTER(OrcPtScene), C.c_float, C.c_float, f32p]
        lib.orc_background_evaluate.argtypes = [C.POINTER(OrcPtScene), f32p, f32p]
        lib._bsdf_ready = True
    return lib


def bsdf_sample(material, in_ts, u0, u1, i0, textures=None, uv=(0.0, 0.0)):
    """Tangent-space BSDF sample: (valid, throughput[3], excident[3], pdf_forw, pdf_back)."""
    lib = _bsdf_lib()
    m = np.ascontiguousarray(material)
    tx = textures or Textures()
    out = np.zeros(9, np.float32)
    lib.orc_bsdf_sample(m.ctypes.data, tx.arr, uv[0], uv[1], _p(np.ascontiguousarray(in_ts, np.float32), f32p), u0, u1,
                        int(i0), _p(out, f32p))
    return bool(out[0]), out[1:4].copy(), out[4:7].copy(), float(out[7]), float(out[8])


def bsdf_evaluate(material, in_ts, ex_ts, textures=None, uv=(0.0, 0.0)):
    """Tangent-space BSDF value[3], pdf_forw, pdf_back."""
    lib = _bsdf_lib()
    m = np.ascontiguousarray(material)
    tx = textures or Textures()
    out = np.zeros(5, np.float32)
    lib.orc_bsdf_evaluate(m.ctypes.data, tx.arr, uv[0], uv[1], _p(np.ascontiguousarray(in_ts, np.float32), f32p),
                          _p(np.ascontiguousarray(ex_ts, np.float32), f32p), _p(out, f32p))
    return out[:3].copy(), float(out[3]), float(out[4])


def bsdf_evaluate_batch(material, in_ts, ex_ts, textures=None, uv=(0.0, 0.0)):
    """bsdf_evaluate over arrays of directions (n, 3): values (n, 3), pdf_forw (n,), pdf_back (n,)."""
    lib = _bsdf_lib()
    m = np.ascontiguousarray(material)
    tx = textures or Textures()
    a = np.ascontiguousarray(in_ts, np.float32).reshape(-1, 3)
    b = np.ascontiguousarray(np.broadcast_to(ex_ts, a.shape), np.float32)
    out = np.zeros((len(a), 5), np.float32)
    lib.orc_bsdf_evaluate_batch(m.ctypes.data, tx.arr, uv[0], uv[1], len(a), _p(a, f32p), _p(b, f32p), _p(out, f32p))
    return out[:, :3], out[:, 3], out[:, 4]


def eval_ndf(ndf, rx, ry, h):
    return _bsdf_lib().orc_eval_ndf(ndf, rx, ry, _p(np.ascontiguousarray(h, np.float32), f32p))


def fresnel_dielectric(n_i, n_t, cos_i):
    ct = np.zeros(1, np.float32)
    f = _bsdf_lib().orc_fresnel_dielectric(n_i, n_t, cos_i, _p(ct, f32p))
    return f, float(ct[0])


def background_scene(background, textures, scene_box, sat=None, flux=(0.0, 0.0, 0.0)):
    """A bare OrcPtScene holding only a background (for the sampler checks)."""
    s = OrcPtScene()
    _fill_background(s, background, textures, sat, flux)
    for k in range(6):
        s.sceneBox[k] = float(scene_box[k])
    return s


def _fill_background(s, background, textures, sat, flux):
    s.textures = textures.arr
    s.numTextures = textures.count
    if background is not None:
        s.background.type = int(background["type"])
        for k in range(3):
            s.background.color[k] = float(background["color"][k])
            s.background.flux[k] = float(flux[k])
        sc = np.broadcast_to(np.asarray(background["scale"], np.float32), (3,))
        for k in range(3):
            s.background.scale[k] = float(sc[k])
        s.background.envmapTexture = int(background["envmapTexture"])
        if sat is not None:
            s.background.sat = _p(sat, f32p)
    else:
        s.background.type = 0
        for k in range(3):
            s.background.scale[k] = 1.0


def background_sample(bscene, u0, u1):
    out = np.zeros(7, np.float32)
    _bsdf_lib().orc_background_sample(C.byref(bscene), u0, u1, _p(out, f32p))
    return out[:3].copy(), float(out[3]), out[4:7].copy()


def background_evaluate(bscene, direction):
    out = np.zeros(4, np.float32)
    _bsdf_lib().orc_background_evaluate(C.byref(bscene), _p(np.ascontiguousarray(direction, np.float32), f32p), _p(out, f32p))
    return out[:3].copy(), float(out[3])


class PtScene:
    """Oracle path-tracer scene over an oracle traversal Scene (single LoD list
    order), with shading arrays concatenated like the geometry, the host-built
    light tree (info, mem, keys, codes) and a pinhole camera dict.
    pos_guide: the "NEE position guide" parameter."""

    def __init__(self, scene, normals, uvs, mat_idx, materials, light_tree, scene_box, camera, textures=None,
                 background=None, background_sat=None, background_flux=(0.0, 0.0, 0.0), pos_guide=False):
        self.scene = scene
        self.textures = textures or Textures()
        self.sat = None if background_sat is None else np.ascontiguousarray(background_sat, np.float32)
        self.normals = np.ascontiguousarray(normals, np.float32)
        self.uvs = np.ascontiguousarray(uvs, np.float32)
        self.mat_idx = np.ascontiguousarray(mat_idx, np.uint16)
        self.materials = np.ascontiguousarray(materials)
        info, mem, keys, codes = light_tree
        self.mem = np.ascontiguousarray(mem if len(mem) else np.zeros(16, np.uint8))
        self.keys = np.ascontiguousarray(keys if len(keys) else np.zeros(1, np.uint64), np.uint64)
        self.codes = np.ascontiguousarray(codes if len(codes) else np.zeros(1, np.uint32), np.uint32)
        s = OrcPtScene()
        s.geo = scene.s
        s.normals = _p(self.normals, f32p)
        s.uvs = _p(self.uvs, f32p)
        s.matIdx = self.mat_idx.ctypes.data
        s.materials = self.materials.ctypes.data
        s.numMaterials = len(self.materials)
        s.lightMem = self.mem.ctypes.data
        s.posRootFlux = float(info["posRootFlux"])
        s.posRootType = int(info["posRootType"])
        s.posInternalCount = int(info["posInternalCount"])
        s.posLightCount = int(info["posLightCount"])
        s.dirRootFlux = float(info["dirRootFlux"])
        s.bgFlux = float(info["backgroundFlux"])
        s.mapKeys = _p(self.keys, u64p)
        s.mapCodes = _p(self.codes, u32p)
        s.mapCount = int(info["mapCount"])
        for k in range(6):
            s.sceneBox[k] = float(scene_box[k])
        for k in range(3):
            s.camPos[k] = float(camera["position"][k])
            s.camDir[k] = float(camera["view_dir"][k])
            s.camUp[k] = float(camera["up"][k])
        s.tanVFov = float(camera["tan_vfov"])
        s.width = int(camera["width"])
        s.height = int(camera["height"])
        _fill_background(s, background, self.textures, self.sat, background_flux)
        s.posGuide = 1 if pos_guide else 0
        self.s = s


def scramble_seed(x):
    lib = load()
    lib.orc_scramble_seed.argtypes = [C.c_uint32]
    lib.orc_scramble_seed.restype = C.c_uint64
    return int(lib.orc_scramble_seed(x))


def pcg_sequence(seed, n):
    """First n outputs of the oracle's PCG64(seed)."""
    lib = load()
    lib.orc_pcg_sequence.argtypes = [C.c_uint32, C.c_int32, u64p]
    out = np.zeros(n, np.uint64)
    lib.orc_pcg_sequence(seed, n, _p(out, u64p))
    return [int(x) for x in out]


def _light_lib():
    lib = load()
    if not hasattr(lib, "_lc_ready"):
        lib.orc_light_connect.argtypes = [C.POINTER(OrcPtScene), C.c_uint64, C.c_uint64, C.c_uint64, f32p, C.c_float,
                                          C.c_float, f32p]
        lib.orc_light_pdf.argtypes = [C.POINTER(OrcPtScene), C.c_int32, C.c_int32, f32p, C.c_float, C.c_float]
        lib.orc_light_pdf.restype = C.c_float
        lib._lc_ready = True
    return lib


def light_connect(pt_scene, index, num_indices, seed, pos, u0=0.25, u1=0.5):
    """One NEE connection through the light tree: (light position (3,),
    creation pdf, diffIrradiance (3,), byte offset of the chosen positional
    light or -1 for the background)."""
    p = np.ascontiguousarray(pos, np.float32)
    out = np.zeros(8, np.float32)
    _light_lib().orc_light_connect(C.byref(pt_scene.s), index, num_indices, seed, _p(p, f32p), u0, u1,
                                   _p(out, f32p))
    return out[:3], float(out[3]), out[4:7], int(out[7])


def light_pdf(pt_scene, inst, prim, ref_pos, surface_params=(0.0, 0.0)):
    """Connect pdf (per area) of hitting emitter (inst, prim) from ref_pos;
    surface_params: the hit's (u, v) on a quad."""
    p = np.ascontiguousarray(ref_pos, np.float32)
    return float(_light_lib().orc_light_pdf(C.byref(pt_scene.s), inst, prim, _p(p, f32p), float(surface_params[0]),
                                             float(surface_params[1])))


def pt_render(pt_scene, seed, min_path_length, max_path_length, nee_count, iterations):
    """Oracle path tracer: running means of Radiance (h,w,3), Hit ID (h,w,2), NaN count."""
    lib = load()
    if not hasattr(lib, "_pt_ready"):
        lib.orc_pt_render.argtypes = [C.POINTER(OrcPtScene), C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                      f32p, f32p, u32p]
        lib.orc_pt_render.restype = C.c_int
        lib._pt_ready = True
    w, h = pt_scene.s.width, pt_scene.s.height
    rad = np.zeros(w * h * 3, np.float32)
    hid = np.zeros(w * h * 2, np.float32)
    nans = C.c_uint32()
    lib.orc_pt_render(C.byref(pt_scene.s), seed, min_path_length, max_path_length, nee_count, iterations,
                      _p(rad, f32p), _p(hid, f32p), C.byref(nans))
    return rad.reshape(h, w, 3), hid.reshape(h, w, 2), nans.value


def pt_render_window(pt_scene, seed, min_path_length, max_path_length, nee_count, iterations, x0, y0, w, h):
    """Oracle Radiance (h, w, 3) of the pixel window at (x0, y0): equal to the
    same pixels of a full render."""
    lib = load()
    lib.orc_pt_render_window.argtypes = [C.POINTER(OrcPtScene)] + [C.c_int32] * 9 + [f32p]
    lib.orc_pt_render_window.restype = C.c_int
    out = np.zeros(w * h * 3, np.float32)
    rc = lib.orc_pt_render_window(C.byref(pt_scene.s), seed, min_path_length, max_path_length, nee_count, iterations,
                                  x0, y0, w, h, _p(out, f32p))
    if rc != 0:
        raise ValueError("window outside the image")
    return out.reshape(h, w, 3)


PATH_STAT_FIELDS = ("ext_rays", "ext_nodes", "ext_tris", "shadow_rays", "shadow_nodes", "shadow_tris",
                    "surface_vertices", "texel_taps", "tree_nodes", "light_records", "paths")


def pt_path_stats(pt_scene, seed, min_path_length, max_path_length, nee_count, x0, y0, w, h):
    """Work counts of one iteration of the window's paths ({field: sum})."""
    lib = load()
    lib.orc_pt_path_stats.argtypes = [C.POINTER(OrcPtScene)] + [C.c_int32] * 8 + [u64p]
    lib.orc_pt_path_stats.restype = C.c_int
    out = np.zeros(11, np.uint64)
    rc = lib.orc_pt_path_stats(C.byref(pt_scene.s), seed, min_path_length, max_path_length, nee_count, x0, y0, w, h,
                               _p(out, u64p))
    if rc != 0:
        raise ValueError("window outside the image")
    return {k: int(v) for k, v in zip(PATH_STAT_FIELDS, out)}


TARGET_CHANNELS = (3, 2, 3, 1, 3, 2, 3, 1)  # Radiance, Hit ID, Position, Depth, Normal, UV, Albedo, Lightness


def pt_render_targets(pt_scene, seed, min_path_length, max_path_length, nee_count, iterations, targets,
                      variance=()):
    """Oracle path tracer: {target: running mean (h, w, channels)} for the
    requested target ids, and the NaN count. With `variance` (target ids, a
    subset of `targets`) it returns ({target: mean}, {target: variance}, nans)."""
    lib = load()
    if not hasattr(lib, "_ptt_ready"):
        lib.orc_pt_render_targets_var.argtypes = [C.POINTER(OrcPtScene), C.c_int32, C.c_int32, C.c_int32,
                                                  C.c_int32, C.c_int32, C.c_uint32, C.POINTER(f32p), C.c_uint32,
                                                  C.POINTER(f32p), u32p]
        lib.orc_pt_render_targets_var.restype = C.c_int
        lib._ptt_ready = True
    w, h = pt_scene.s.width, pt_scene.s.height
    mask = vmask = 0
    bufs, vbufs = {}, {}
    ptrs = (f32p * 8)()
    vptrs = (f32p * 8)()
    for t in targets:
        mask |= 1 << t
        bufs[t] = np.zeros(w * h * TARGET_CHANNELS[t], np.float32)
        ptrs[t] = _p(bufs[t], f32p)
    for t in variance:
        if t not in bufs:
            raise ValueError("variance of target %d requested without the target" % t)
        vmask |= 1 << t
        vbufs[t] = np.zeros(w * h * TARGET_CHANNELS[t], np.float32)
        vptrs[t] = _p(vbufs[t], f32p)
    nans = C.c_uint32()
    lib.orc_pt_render_targets_var(C.byref(pt_scene.s), seed, min_path_length, max_path_length, nee_count,
                                  iterations, mask, ptrs, vmask, vptrs, C.byref(nans))
    means = {t: b.reshape(h, w, TARGET_CHANNELS[t]) for t, b in bufs.items()}
    if variance:
        return means, {t: b.reshape(h, w, TARGET_CHANNELS[t]) for t, b in vbufs.items()}, nans.value
    return means, nans.value


WIREFRAME_RNG_START = 0x2f0ae9bc6431af73  # PCG64() (rng.hpp:72-74)


def wireframe_render(pt_scene, iterations, line_width=1.0, max_trace_depth=1000, rng_state=WIREFRAME_RNG_START,
                     variance=False):
    """Oracle wireframe renderer: (Border mean (h, w, 3), Border variance or
    None, the renderer generator's state afterwards)."""
    lib = load()
    if not hasattr(lib, "_wire_ready"):
        lib.orc_wireframe_render.argtypes = [C.POINTER(OrcPtScene), C.c_int32, C.c_float, C.c_int32, u64p, f32p,
                                             f32p]
        lib.orc_wireframe_render.restype = C.c_int
        lib._wire_ready = True
    w, h = pt_scene.s.width, pt_scene.s.height
    border = np.zeros(w * h * 3, np.float32)
    var = np.zeros(w * h * 3, np.float32) if variance else None
    state = np.array([rng_state], np.uint64)
    lib.orc_wireframe_render(C.byref(pt_scene.s), iterations, line_width, max_trace_depth, _p(state, u64p),
                             _p(border, f32p), _p(var, f32p) if variance else None)
    return border.reshape(h, w, 3), (var.reshape(h, w, 3) if variance else None), int(state[0])


def lt_render(pt_scene, seed, min_path_length, max_path_length, iterations, density=False, variance=False):
    """Oracle light tracer: (means, variances or None), each (h, w, 5) with
    Radiance (3), Lightness and Density per pixel."""
    lib = load()
    if not hasattr(lib, "_lt_ready"):
        lib.orc_lt_render.argtypes = [C.POINTER(OrcPtScene), C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                      f32p, f32p]
        lib.orc_lt_render.restype = C.c_int
        lib._lt_ready = True
    w, h = pt_scene.s.width, pt_scene.s.height
    mean = np.zeros(w * h * 5, np.float32)
    var = np.zeros(w * h * 5, np.float32) if variance else None
    lib.orc_lt_render(C.byref(pt_scene.s), seed, min_path_length, max_path_length, iterations, 1 if density else 0,
                      _p(mean, f32p), _p(var, f32p) if variance else None)
    return mean.reshape(h, w, 5), (var.reshape(h, w, 5) if variance else None)


def pt_render_replay(base_scene, eval_scene, seed, min_path_length, max_path_length, nee_count, iterations):
    """Radiance (h, w, 3) of eval_scene rendered along the BSDF directions the
    same paths sampled in base_scene (the estimator whose roughness
    derivative the adjoint computes)."""
    lib = load()
    if not hasattr(lib, "_replay_ready"):
        lib.orc_pt_render_replay.argtypes = [C.POINTER(OrcPtScene), C.POINTER(OrcPtScene), C.c_int32, C.c_int32,
                                             C.c_int32, C.c_int32, C.c_int32, f32p]
        lib.orc_pt_render_replay.restype = C.c_int
        lib._replay_ready = True
    w, h = base_scene.s.width, base_scene.s.height
    out = np.zeros(w * h * 3, np.float32)
    lib.orc_pt_render_replay(C.byref(base_scene.s), C.byref(eval_scene.s), seed, min_path_length, max_path_length,
                             nee_count, iterations, _p(out, f32p))
    return out.reshape(h, w, 3)


def pt_adjoint(pt_scene, seed, min_path_length, max_path_length, nee_count, iterations, delta, tex_shapes,
               fixed_decisions=False):
    """FP64 adjoint of the oracle: delta (h, w, 3) = dLoss/dRadiance of the
    image of `iterations` iterations from a reset. tex_shapes: {texture id:
    (h, w)} of the albedo / roughness textures to differentiate. The
    derivative of the expected image, or with fixed_decisions the derivative
    of this realization (every random decision and direction held). Returns
    ({id: (h, w, 4) float64}, emission gradient (materials, 3) float64)."""
    lib = load()
    if not hasattr(lib, "_adj_ready"):
        dp = C.POINTER(C.c_double)
        lib.orc_pt_adjoint.argtypes = [C.POINTER(OrcPtScene), C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                       f32p, C.POINTER(dp), dp, C.c_int32]
        lib.orc_pt_adjoint.restype = C.c_int
        lib._adj_ready = True
    dp = C.POINTER(C.c_double)
    n_tex = max(pt_scene.textures.count, 1)
    grads = {}
    ptrs = (dp * n_tex)()
    for t, (h, w) in tex_shapes.items():
        grads[t] = np.zeros((h, w, 4), np.float64)
        ptrs[t] = grads[t].ctypes.data_as(dp)
    em = np.zeros((len(pt_scene.materials), 3), np.float64)
    d = np.ascontiguousarray(delta, np.float32)
    lib.orc_pt_adjoint(C.byref(pt_scene.s), seed, min_path_length, max_path_length, nee_count, iterations,
                       _p(d, f32p), ptrs, em.ctypes.data_as(dp), 1 if fixed_decisions else 0)
    return grads, em


def set_mutation(m):
    """Switches on one planted estimator mistake of the float oracle (0 = none)
    for the closed-form estimator tests; returns the previous setting."""
    lib = load()
    lib.orc_set_mutation.argtypes = [C.c_int32]
    lib.orc_set_mutation.restype = C.c_int
    return lib.orc_set_mutation(int(m))


class Fp64Scene(C.Structure):
    _fields_ = [
        ("verts", f32p), ("numVerts", C.c_int32), ("tris", u32p), ("numTris", C.c_int32), ("normals", f32p),
        ("matIdx", C.c_void_p), ("matType", i32p), ("albedo", f32p), ("emission", f32p),
        ("numMaterials", C.c_int32), ("pointLights", f32p), ("numPointLights", C.c_int32),
        ("background", C.c_float * 3), ("camPos", C.c_float * 3), ("camDir", C.c_float * 3),
        ("camUp", C.c_float * 3), ("tanVFov", C.c_float), ("width", C.c_int32), ("height", C.c_int32),
    ]


def fp64_render(verts, normals, tris, mat_idx, materials, camera, max_path_length, spp, seed=1,
                point_lights=None, background=(0.0, 0.0, 0.0)):
    """Independent FP64 estimator (pt_fp64.cpp) of the path tracer's pixel
    integral for Lambert / emissive scenes with point lights (records with
    'position' and 'intensity') and a uniform background radiance. Returns the
    per-pixel mean (h, w, 3) and the variance of one sample (h, w, 3), float64."""
    lib = load()
    dp = C.POINTER(C.c_double)
    lib.orc64_render.argtypes = [C.POINTER(Fp64Scene), C.c_int32, C.c_int32, C.c_uint64, dp, dp]
    lib.orc64_render.restype = C.c_int
    mats = np.asarray(materials)
    if (mats["albedoTexture"] >= 0).any():
        raise ValueError("the FP64 estimator takes constant albedos only")
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a

    mt = arr(mats["type"], np.int32)
    if not np.isin(mt, (0, 1)).all():
        raise ValueError("the FP64 estimator takes Lambert and emissive materials only")
    alb = arr(mats["albedo"], np.float32)
    em = arr(mats["radiance"].astype(np.float64) * mats["scale"].astype(np.float64), np.float32)
    v = arr(verts, np.float32)
    n = arr(normals, np.float32)
    t = arr(tris, np.uint32)
    mi = arr(mat_idx, np.uint16)
    pl = np.zeros((0, 6), np.float32)
    if point_lights is not None and len(point_lights):
        pl = np.concatenate([np.asarray(point_lights["position"], np.float32),
                             np.asarray(point_lights["intensity"], np.float32)], 1)
    pl = arr(pl.reshape(-1) if len(pl) else np.zeros(6, np.float32), np.float32)
    s = Fp64Scene()
    s.verts, s.numVerts = _p(v, f32p), len(v)
    s.tris, s.numTris = _p(t, u32p), len(t)
    s.normals = _p(n, f32p)
    s.matIdx = mi.ctypes.data
    s.matType, s.albedo, s.emission, s.numMaterials = _p(mt, i32p), _p(alb, f32p), _p(em, f32p), len(mats)
    s.pointLights = _p(pl, f32p)
    s.numPointLights = 0 if point_lights is None else len(point_lights)
    for k in range(3):
        s.background[k] = float(background[k])
        s.camPos[k] = float(camera["position"][k])
        s.camDir[k] = float(camera["view_dir"][k])
        s.camUp[k] = float(camera["up"][k])
    s.tanVFov = float(camera["tan_vfov"])
    s.width, s.height = int(camera["width"]), int(camera["height"])
    w, h = s.width, s.height
    mean = np.zeros(w * h * 3, np.float64)
    var = np.zeros(w * h * 3, np.float64)
    rc = lib.orc64_render(C.byref(s), max_path_length, spp, seed, mean.ctypes.data_as(dp), var.ctypes.data_as(dp))
    if rc != 0:
        raise RuntimeError("orc64_render failed (%d)" % rc)
    return mean.reshape(h, w, 3), var.reshape(h, w, 3)
