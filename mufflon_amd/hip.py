"""ctypes binding of the HIP library (include/mufflon_hip.h).

The library is loaded from mufflon_amd/lib/libmufflon_hip.so, built in-tree
by mufflon_amd.build. There is no fallback: if the library is missing or
fails to load, the import of the device raises.
"""
import ctypes as C
import os

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libmufflon_hip.so")
_lib = None

i32 = C.c_int32
u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)
u64p = C.POINTER(C.c_uint64)
vp = C.c_void_p

# C-ABI record layouts (include/mufflon_hip.h)
MATERIAL_DTYPE = np.dtype([("type", "<i4"), ("albedo", "<f4", 3), ("radiance", "<f4", 3), ("scale", "<f4", 3),
                           ("roughness", "<f4", 2), ("ndf", "<i4"), ("shadowing", "<i4"),
                           ("refractionIndex", "<f4"), ("absorption", "<f4", 3), ("outerRefractionIndex", "<f4"),
                           ("outerAbsorption", "<f4", 3), ("albedoTexture", "<i4"), ("roughnessTexture", "<i4"),
                           ("emissionTexture", "<i4"), ("reserved", "<i4", 7)])
assert MATERIAL_DTYPE.itemsize == 128
BACKGROUND_DTYPE = np.dtype([("type", "<i4"), ("color", "<f4", 3), ("scale", "<f4", 3), ("envmapTexture", "<i4")])
BACKGROUND_COLORED = 0
BACKGROUND_ENVMAP = 1
SAMPLING_NEAREST = 0
SAMPLING_LINEAR = 1
PINHOLE_DTYPE = np.dtype([("position", "<f4", 3), ("tanVFov", "<f4"), ("viewDir", "<f4", 3), ("up", "<f4", 3),
                          ("width", "<i4"), ("height", "<i4")])
PT_PARAMS_DTYPE = np.dtype([("seed", "<i4"), ("minPathLength", "<i4"), ("maxPathLength", "<i4"),
                            ("neeCount", "<i4"), ("neePositionGuide", "<i4")])
POINT_LIGHT_DTYPE = np.dtype([("position", "<f4", 3), ("intensity", "<f4", 3)])
AREA_LIGHT_DTYPE = np.dtype([("points", "<f4", 9), ("uv", "<f4", 6), ("material", "<i4"), ("instance", "<i4"),
                             ("primitive", "<i4")])
AREA_LIGHT_EX_DTYPE = np.dtype([("kind", "<i4"), ("points", "<f4", 12), ("uv", "<f4", 8), ("radius", "<f4"),
                                ("material", "<i4"), ("instance", "<i4"), ("primitive", "<i4")])
AREA_TRIANGLE, AREA_QUAD, AREA_SPHERE = 0, 1, 2
LIGHT_TREE_INFO_DTYPE = np.dtype([("posRootFlux", "<f4"), ("posRootType", "<u4"), ("posInternalCount", "<u4"),
                                  ("posLightCount", "<u4"), ("memBytes", "<u4"), ("mapCount", "<i4"),
                                  ("dirRootFlux", "<f4"), ("backgroundFlux", "<f4")])
ADJOINT_EXPECTED = 0  # mfh_pt_set_adjoint_mode
ADJOINT_FIXED_DECISIONS = 1
TARGET_RADIANCE = 0
TARGET_HIT_ID = 1
TARGET_POSITION = 2
TARGET_DEPTH = 3
TARGET_NORMAL = 4
TARGET_UV = 5
TARGET_ALBEDO = 6
TARGET_LIGHTNESS = 7
TARGET_CHANNELS = (3, 2, 3, 1, 3, 2, 3, 1)
LT_RADIANCE, LT_LIGHTNESS, LT_DENSITY = 0, 1, 2
PT_MEGAKERNEL = 0
PT_WAVEFRONT = 1
TUNING_DTYPE = np.dtype([("traceBlock", "<i4"), ("stackDepth", "<i4"), ("queueChunk", "<i4"), ("sortBounces", "<i4"),
                         ("refillBudget", "<i4"), ("leafBatch", "<i4"), ("tailAfter", "<i4")])

# name -> argtypes; every function returns int status
_SIGS = {
    "mfh_create": [i32, C.POINTER(vp)],
    "mfh_destroy": [vp],
    "mfh_synchronize": [vp],
    "mfh_alloc": [vp, C.c_size_t, C.POINTER(vp)],
    "mfh_free": [vp, vp],
    "mfh_upload": [vp, vp, vp, C.c_size_t],
    "mfh_download": [vp, vp, vp, C.c_size_t],
    "mfh_lod_create": [vp, f32p, i32, u32p, i32, i32p],
    "mfh_lod_create_mixed": [vp, f32p, i32, u32p, i32, u32p, i32, f32p, i32, i32p],
    "mfh_lbvh_build": [vp, i32, f32p],
    "mfh_lbvh_get": [vp, i32, i32p, vp, i32p],
    "mfh_lbvh_build_stages": [vp, i32, f32p, u64p, u64p, i32p, i32p, f32p, i32p],
    "mfh_lbvh_build_noalloc": [vp, vp, i32, vp, i32, f32p, i32p, vp, vp],
    "mfh_scene_create": [vp, i32p, f32p, i32, i32p],
    "mfh_scene_get": [vp, i32, f32p, i32p, vp, i32p],
    "mfh_trace_closest": [vp, i32, vp, i32, vp, vp],
    "mfh_trace_closest_stats": [vp, i32, vp, i32, vp, vp, vp],
    "mfh_trace_any": [vp, i32, vp, i32, vp],
    "mfh_lod_set_shading": [vp, i32, f32p, f32p, vp],
    "mfh_pt_create": [vp, i32, vp, i32, vp, i32, vp, vp, i32p],
    "mfh_pt_reset": [vp, i32],
    "mfh_pt_set_params": [vp, i32, vp],
    "mfh_pt_set_camera": [vp, i32, vp],
    "mfh_pt_iterate": [vp, i32, i32],
    "mfh_pt_wireframe_reset": [vp, i32, i32],
    "mfh_pt_wireframe_iterate": [vp, i32, i32, C.c_float, i32],
    "mfh_pt_wireframe_get": [vp, i32, i32, vp],
    "mfh_pt_lighttracer_reset": [vp, i32, i32],
    "mfh_pt_lighttracer_iterate": [vp, i32, i32, i32],
    "mfh_pt_lighttracer_get": [vp, i32, i32, i32, vp],
    "mfh_pt_get_target": [vp, i32, i32, f32p],
    "mfh_pt_enable_target": [vp, i32, i32, i32],
    "mfh_pt_enable_variance": [vp, i32, i32, i32],
    "mfh_pt_get_variance": [vp, i32, i32, f32p],
    "mfh_pt_set_mode": [vp, i32, i32, i32],
    "mfh_pt_adjoint": [vp, i32, i32, f32p],
    "mfh_set_traversal_leaf_limit": [vp, i32],
    "mfh_pt_set_adjoint_mode": [vp, i32, i32],
    "mfh_batch_create": [vp, i32p, i32, i32, i32p],
    "mfh_batch_reset": [vp, i32],
    "mfh_batch_iterate": [vp, i32, i32],
    "mfh_batch_set_tuning": [vp, i32, vp],
    "mfh_batch_autotune": [vp, i32, i32, vp, f32p],
    "mfh_pt_get_texture_gradient": [vp, i32, i32, f32p],
    "mfh_pt_get_emission_gradient": [vp, i32, i32, f32p],
    "mfh_pt_set_material": [vp, i32, i32, vp],
    "mfh_texture_update": [vp, i32, f32p],
    "mfh_pt_set_tuning": [vp, i32, vp],
    "mfh_pt_get_tuning": [vp, i32, vp],
    "mfh_pt_autotune": [vp, i32, i32, vp, f32p],
    "mfh_pt_get_iteration": [vp, i32, i32p],
    "mfh_pt_get_nan_count": [vp, i32, u32p],
    "mfh_pt_get_light_tree": [vp, i32, vp, vp, u64p, u32p],
    "mfh_light_tree_build": [vp, i32, vp, i32, f32p, vp, i32, vp, vp, C.c_size_t, u64p, u32p, i32],
    "mfh_light_tree_build_ex": [vp, i32, vp, i32, f32p, vp, i32, vp, vp, C.c_size_t, u64p, u32p, i32],
    "mfh_texture_create": [vp, i32, i32, i32, i32, f32p, i32p],
    "mfh_background_build": [vp, f32p, i32, i32, i32, f32p, f32p, f32p],
    "mfh_pt_set_background": [vp, i32, vp],
    "mfh_timer_begin": [vp],
    "mfh_timer_end": [vp, f32p],
}


def library_path():
    return _LIB_PATH


def load():
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError("HIP library not built: %s (run __graft_entry__.build())" % _LIB_PATH)
    lib = C.CDLL(_LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.mfh_last_error.argtypes = []
    lib.mfh_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def camera_record(cam):
    r = np.zeros(1, PINHOLE_DTYPE)
    r["position"] = cam["position"]
    r["tanVFov"] = cam["tan_vfov"]
    r["viewDir"] = cam["view_dir"]
    r["up"] = cam["up"]
    r["width"] = cam["width"]
    r["height"] = cam["height"]
    return r


def params_record(seed=0, min_path_length=0, max_path_length=16, nee_count=1, nee_position_guide=False):
    r = np.zeros(1, PT_PARAMS_DTYPE)
    r["seed"], r["minPathLength"], r["maxPathLength"], r["neeCount"] = seed, min_path_length, max_path_length, \
        nee_count
    r["neePositionGuide"] = 1 if nee_position_guide else 0
    return r


def background_record(kind=BACKGROUND_COLORED, color=(0.0, 0.0, 0.0), scale=1.0, envmap_texture=-1):
    r = np.zeros(1, BACKGROUND_DTYPE)
    r["type"] = kind
    r["color"] = color
    r["scale"] = scale
    r["envmapTexture"] = envmap_texture
    return r


def background_build(background, env_texels, scene_box):
    """Host-only background preparation: (flux[3], sat (h, w) or None).
    env_texels: (h, w, c) float32 for an ENVMAP background, else None."""
    lib = load()
    bg = background if isinstance(background, np.ndarray) else background_record(**background)
    box = np.ascontiguousarray(scene_box, np.float32)
    flux = np.zeros(3, np.float32)
    if int(bg["type"][0]) == BACKGROUND_ENVMAP:
        t = np.ascontiguousarray(env_texels, np.float32)
        if t.ndim == 2:
            t = t[:, :, None]
        h, w, c = t.shape
        sat = np.zeros((h, w), np.float32)
        rc = lib.mfh_background_build(bg.ctypes.data, _ptr(t, f32p), w, h, c, _ptr(box, f32p), _ptr(flux, f32p),
                                      _ptr(sat, f32p))
    else:
        sat = None
        rc = lib.mfh_background_build(bg.ctypes.data, None, 0, 0, 0, _ptr(box, f32p), _ptr(flux, f32p), None)
    if rc != 0:
        raise HipError(rc, lib.mfh_last_error().decode())
    return flux, sat


def light_tree_build(area_lights, point_lights, scene_box, materials):
    """Host-only light-tree build (no device): returns (info, mem, keys, codes).
    area_lights: AREA_LIGHT_DTYPE (triangles) or AREA_LIGHT_EX_DTYPE records."""
    lib = load()
    ex = area_lights is not None and getattr(area_lights, "dtype", None) == AREA_LIGHT_EX_DTYPE
    dt = AREA_LIGHT_EX_DTYPE if ex else AREA_LIGHT_DTYPE
    area = np.ascontiguousarray(area_lights, dt) if area_lights is not None else np.zeros(0, dt)
    pts = np.ascontiguousarray(point_lights, POINT_LIGHT_DTYPE) if point_lights is not None else \
        np.zeros(0, POINT_LIGHT_DTYPE)
    mats = np.ascontiguousarray(materials, MATERIAL_DTYPE)
    box = np.ascontiguousarray(scene_box, np.float32)
    n = len(area) + len(pts)
    info = np.zeros(1, LIGHT_TREE_INFO_DTYPE)
    mem = np.zeros(max(192 * n, 16), np.uint8)
    keys = np.zeros(max(len(area), 1), np.uint64)
    codes = np.zeros(max(len(area), 1), np.uint32)
    fn = lib.mfh_light_tree_build_ex if ex else lib.mfh_light_tree_build
    rc = fn(area.ctypes.data if len(area) else None, len(area),
                                  pts.ctypes.data if len(pts) else None, len(pts), _ptr(box, f32p), mats.ctypes.data,
                                  len(mats), info.ctypes.data, mem.ctypes.data, mem.nbytes, _ptr(keys, u64p),
                                  _ptr(codes, u32p), len(keys))
    if rc != 0:
        raise HipError(rc, lib.mfh_last_error().decode())
    i = info[0]
    return i, mem[: i["memBytes"]].copy(), keys[: i["mapCount"]].copy(), codes[: i["mapCount"]].copy()


class HipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("mfh error %d: %s" % (code, msg))
        self.code = code


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


class DeviceBuffer:
    """Device allocation owned by a context (freed explicitly or with it)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = vp()
        ctx._call("mfh_alloc", ctx.handle, self.nbytes, C.byref(p))
        self.ptr = p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.ctx._call("mfh_upload", self.ctx.handle, self.ptr, arr.ctypes.data, arr.nbytes)

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        self.ctx._call("mfh_download", self.ctx.handle, out.ctypes.data, self.ptr, out.nbytes)
        return out

    def free(self):
        if self.ptr:
            self.ctx._call("mfh_free", self.ctx.handle, self.ptr)
            self.ptr = None


class HipContext:
    """One device context (mfh_create); the Python face of the C-ABI."""

    def __init__(self, device=0):
        self.lib = load()
        h = vp()
        self._call("mfh_create", int(device), C.byref(h))
        self.handle = h.value

    def _call(self, name, *args):
        rc = getattr(self.lib, name)(*args)
        if rc != 0:
            raise HipError(rc, self.lib.mfh_last_error().decode())
        return rc

    def close(self):
        if getattr(self, "handle", None):
            self._call("mfh_destroy", self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self):
        self._call("mfh_synchronize", self.handle)

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        b = DeviceBuffer(self, max(arr.nbytes, 1))
        if arr.nbytes:
            b.upload(arr)
        return b

    # geometry / BVH -------------------------------------------------------
    def lod_create(self, vertices, tris):
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
        out = i32()
        self._call("mfh_lod_create", self.handle, _ptr(v, f32p), len(v), _ptr(t, u32p), len(t), C.byref(out))
        return out.value

    def lod_create_mixed(self, vertices, tris=None, quads=None, spheres=None):
        """LoD of triangles (n x 3 vertex indices), quads (n x 4: p00, p10, p11,
        p01) and spheres (n x 4: center xyz, radius); ids run in that order."""
        v = np.ascontiguousarray(vertices if vertices is not None else np.zeros((0, 3)), np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(tris if tris is not None else np.zeros((0, 3)), np.uint32).reshape(-1, 3)
        q = np.ascontiguousarray(quads if quads is not None else np.zeros((0, 4)), np.uint32).reshape(-1, 4)
        sp = np.ascontiguousarray(spheres if spheres is not None else np.zeros((0, 4)), np.float32).reshape(-1, 4)
        out = i32()
        self._call("mfh_lod_create_mixed", self.handle, _ptr(v, f32p) if len(v) else None, len(v),
                   _ptr(t, u32p) if len(t) else None, len(t), _ptr(q, u32p) if len(q) else None, len(q),
                   _ptr(sp, f32p) if len(sp) else None, len(sp), C.byref(out))
        return out.value

    def lbvh_build(self, lod, bbox=None):
        b = None if bbox is None else np.ascontiguousarray(bbox, dtype=np.float32)
        self._call("mfh_lbvh_build", self.handle, lod, _ptr(b, f32p))

    def set_traversal_leaf_limit(self, limit):
        """Longest traversal leaf of later LoD builds (0: the reference BVH)."""
        self._call("mfh_set_traversal_leaf_limit", self.handle, int(limit))

    def lbvh_get(self, lod, num_prims):
        n = i32()
        self._call("mfh_lbvh_get", self.handle, lod, C.byref(n), None, None)
        nodes = np.zeros(max(n.value, 0) * 64, dtype=np.uint8)
        ids = np.zeros(num_prims, dtype=np.int32)
        self._call("mfh_lbvh_get", self.handle, lod, C.byref(n), nodes.ctypes.data if len(nodes) else None,
                   _ptr(ids, i32p))
        return n.value, nodes, ids

    def lbvh_build_noalloc(self, d_verts, num_verts, d_tris, num_tris, bbox, d_nodes, d_prim_ids):
        """LBVH of a LoD in device buffers without any device allocation
        (mfh_lbvh_build_noalloc); returns the number of internal nodes."""
        b = np.ascontiguousarray(bbox, dtype=np.float32)
        n = i32()
        self._call("mfh_lbvh_build_noalloc", self.handle, d_verts, int(num_verts), d_tris, int(num_tris),
                   _ptr(b, f32p), C.byref(n), d_nodes, d_prim_ids)
        return n.value

    def lbvh_build_stages(self, lod, num_prims, bbox=None):
        n = num_prims
        b = None if bbox is None else np.ascontiguousarray(bbox, dtype=np.float32)
        uk = np.zeros(n, np.uint64)
        sk = np.zeros(n, np.uint64)
        ids = np.zeros(n, np.int32)
        par = np.zeros(max(2 * n - 1, 1), np.int32)
        boxes = np.zeros(max(2 * n - 1, 1) * 8, np.float32)
        offs = np.zeros(max(n - 1, 1), np.int32)
        self._call("mfh_lbvh_build_stages", self.handle, lod, _ptr(b, f32p), _ptr(uk, u64p), _ptr(sk, u64p),
                   _ptr(ids, i32p), _ptr(par, i32p), _ptr(boxes, f32p), _ptr(offs, i32p))
        return dict(unsorted_keys=uk, sorted_keys=sk, sorted_ids=ids, parents=par, boxes=boxes, offsets=offs)

    def scene_create(self, inst_lod, world_to_inst):
        il = np.ascontiguousarray(inst_lod, dtype=np.int32)
        w = np.ascontiguousarray(world_to_inst, dtype=np.float32).reshape(-1, 12)
        assert len(il) == len(w)
        out = i32()
        self._call("mfh_scene_create", self.handle, _ptr(il, i32p), _ptr(w, f32p), len(il), C.byref(out))
        return out.value

    def scene_get(self, scene, num_instances):
        box = np.zeros(6, np.float32)
        n = i32()
        self._call("mfh_scene_get", self.handle, scene, _ptr(box, f32p), C.byref(n), None, None)
        nodes = np.zeros(max(n.value, 0) * 64, np.uint8)
        ids = np.zeros(num_instances, np.int32)
        self._call("mfh_scene_get", self.handle, scene, _ptr(box, f32p), C.byref(n),
                   nodes.ctypes.data if len(nodes) else None, _ptr(ids, i32p))
        return box, n.value, nodes, ids

    # tracing ----------------------------------------------------------------
    def trace_closest_dev(self, scene, d_rays, n, d_hits, d_bary=None):
        self._call("mfh_trace_closest", self.handle, scene, d_rays.ptr, int(n), d_hits.ptr,
                   d_bary.ptr if d_bary is not None else None)

    def trace_closest_stats_dev(self, scene, d_rays, n, d_hits, d_bary, d_stats):
        self._call("mfh_trace_closest_stats", self.handle, scene, d_rays.ptr, int(n), d_hits.ptr,
                   d_bary.ptr if d_bary is not None else None, d_stats.ptr)

    def trace_stats(self, scene, rays):
        """Per-ray (internal nodes visited, triangles tested) of closest hit."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = len(rays)
        dr = self.to_device(rays)
        dh = self.alloc(16 * max(n, 1))
        ds = self.alloc(8 * max(n, 1))
        try:
            self.trace_closest_stats_dev(scene, dr, n, dh, None, ds)
            self.synchronize()
            st = ds.download(np.uint32, 2 * n).reshape(n, 2)
        finally:
            dr.free(); dh.free(); ds.free()
        return st[:, 0].copy(), st[:, 1].copy()

    def trace_any_dev(self, scene, d_rays, n, d_occl):
        self._call("mfh_trace_any", self.handle, scene, d_rays.ptr, int(n), d_occl.ptr)

    def trace_closest(self, scene, rays):
        """Host convenience: rays (N,8) float32 -> t, inst, prim, bary, status."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = len(rays)
        dr = self.to_device(rays)
        dh = self.alloc(16 * max(n, 1))
        db = self.alloc(8 * max(n, 1))
        try:
            self.trace_closest_dev(scene, dr, n, dh, db)
            self.synchronize()
            h = dh.download(np.uint32, 4 * n).reshape(n, 4)
            bary = db.download(np.float32, 2 * n).reshape(n, 2)
        finally:
            dr.free(); dh.free(); db.free()
        t = h[:, 0].view(np.float32).copy()
        return t, h[:, 1].view(np.int32).copy(), h[:, 2].view(np.int32).copy(), bary, h[:, 3].view(np.float32).copy()

    def trace_any(self, scene, rays):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = len(rays)
        dr = self.to_device(rays)
        do = self.alloc(4 * max(n, 1))
        try:
            self.trace_any_dev(scene, dr, n, do)
            self.synchronize()
            o = do.download(np.int32, n)
        finally:
            dr.free(); do.free()
        return o

    # path tracer --------------------------------------------------------------
    def lod_set_shading(self, lod, normals, uvs, mat_idx):
        n = np.ascontiguousarray(normals, np.float32)
        uv = np.ascontiguousarray(uvs, np.float32)
        m = np.ascontiguousarray(mat_idx, np.uint16)
        self._call("mfh_lod_set_shading", self.handle, lod, _ptr(n, f32p) if n.size else None,
                   _ptr(uv, f32p) if uv.size else None, m.ctypes.data)

    def texture_create(self, texels, sampling=SAMPLING_LINEAR):
        """Texture from (h, w) or (h, w, c) float32 texels; returns its id."""
        t = np.ascontiguousarray(texels, np.float32)
        if t.ndim == 2:
            t = t[:, :, None]
        h, w, c = t.shape
        out = i32()
        self._call("mfh_texture_create", self.handle, w, h, c, sampling, _ptr(t, f32p), C.byref(out))
        return out.value

    def pt_set_background(self, pt, background):
        bg = background if isinstance(background, np.ndarray) else background_record(**background)
        self._call("mfh_pt_set_background", self.handle, pt, bg.ctypes.data)

    def pt_create(self, scene, materials, camera, params, point_lights=None):
        mats = np.ascontiguousarray(materials, MATERIAL_DTYPE)
        cam = camera if isinstance(camera, np.ndarray) else camera_record(camera)
        par = params if isinstance(params, np.ndarray) else params_record(**params)
        pts = np.ascontiguousarray(point_lights, POINT_LIGHT_DTYPE) if point_lights is not None else None
        out = i32()
        self._call("mfh_pt_create", self.handle, scene, mats.ctypes.data, len(mats),
                   pts.ctypes.data if pts is not None and len(pts) else None, 0 if pts is None else len(pts),
                   cam.ctypes.data, par.ctypes.data, C.byref(out))
        self._pt_size = getattr(self, "_pt_size", {})
        self._pt_size[out.value] = (int(cam["width"][0]), int(cam["height"][0]))
        return out.value

    def pt_reset(self, pt):
        self._call("mfh_pt_reset", self.handle, pt)

    def pt_set_params(self, pt, params):
        par = params if isinstance(params, np.ndarray) else params_record(**params)
        self._call("mfh_pt_set_params", self.handle, pt, par.ctypes.data)

    def pt_set_camera(self, pt, camera):
        cam = camera if isinstance(camera, np.ndarray) else camera_record(camera)
        self._call("mfh_pt_set_camera", self.handle, pt, cam.ctypes.data)
        self._pt_size[pt] = (int(cam["width"][0]), int(cam["height"][0]))

    def pt_iterate(self, pt, iterations=1):
        self._call("mfh_pt_iterate", self.handle, pt, int(iterations))

    def wireframe_reset(self, pt, variance=False):
        self._call("mfh_pt_wireframe_reset", self.handle, pt, 1 if variance else 0)

    def wireframe_iterate(self, pt, iterations=1, line_width=1.0, max_trace_depth=1000):
        self._call("mfh_pt_wireframe_iterate", self.handle, pt, int(iterations), float(line_width),
                   int(max_trace_depth))

    def wireframe_border(self, pt, variance=False):
        """The Border target (height, width, 3) of the wireframe renderer."""
        w, h = self._pt_size[pt]
        out = np.empty(w * h * 3, np.float32)
        self._call("mfh_pt_wireframe_get", self.handle, pt, 1 if variance else 0, out.ctypes.data)
        return out.reshape(h, w, 3)


    # light tracer over a path tracer's scene, camera and lights
    def lighttracer_reset(self, pt, variance=False):
        self._call("mfh_pt_lighttracer_reset", self.handle, pt, 1 if variance else 0)

    def lighttracer_iterate(self, pt, iterations=1, density=False):
        self._call("mfh_pt_lighttracer_iterate", self.handle, pt, iterations, 1 if density else 0)

    def lighttracer_target(self, pt, target=LT_RADIANCE, variance=False):
        """Running mean (or variance) of a light tracer target, (h, w, channels)."""
        w, h = self._pt_size[pt]
        nc = 3 if target == LT_RADIANCE else 1
        out = np.zeros(w * h * nc, np.float32)
        self._call("mfh_pt_lighttracer_get", self.handle, pt, target, 1 if variance else 0, out.ctypes.data)
        return out.reshape(h, w, nc)
    def batch_create(self, pts, use_graph=False):
        ids = np.ascontiguousarray(pts, np.int32)
        out = i32()
        self._call("mfh_batch_create", self.handle, _ptr(ids, i32p), len(ids), 1 if use_graph else 0, C.byref(out))
        return out.value

    def batch_reset(self, batch):
        self._call("mfh_batch_reset", self.handle, batch)

    def batch_iterate(self, batch, iterations=1):
        self._call("mfh_batch_iterate", self.handle, batch, iterations)

    def batch_set_tuning(self, batch, trace_block=256, stack_depth=16, queue_chunk=256, refill_budget=0,
                         leaf_batch=1, tail_after=0):
        r = np.zeros(1, TUNING_DTYPE)
        r["traceBlock"], r["stackDepth"], r["queueChunk"], r["sortBounces"] = trace_block, stack_depth, queue_chunk, 0
        r["refillBudget"] = refill_budget
        r["leafBatch"] = leaf_batch
        r["tailAfter"] = tail_after
        self._call("mfh_batch_set_tuning", self.handle, batch, r.ctypes.data)

    def batch_autotune(self, batch, iterations=1):
        r = np.zeros(1, TUNING_DTYPE)
        ms = np.zeros(1, np.float32)
        self._call("mfh_batch_autotune", self.handle, batch, iterations, r.ctypes.data, _ptr(ms, f32p))
        return {k: int(r[k][0]) for k in TUNING_DTYPE.names}, float(ms[0])

    def texture_update(self, tex, texels):
        t = np.ascontiguousarray(texels, np.float32)
        self._call("mfh_texture_update", self.handle, tex, _ptr(t, f32p))

    def pt_set_material(self, pt, index, material):
        m = np.ascontiguousarray(material, MATERIAL_DTYPE).reshape(1)
        self._call("mfh_pt_set_material", self.handle, pt, index, m.ctypes.data)

    def pt_adjoint(self, pt, iterations, radiance_adjoint):
        d = np.ascontiguousarray(radiance_adjoint, np.float32)
        self._call("mfh_pt_adjoint", self.handle, pt, iterations, _ptr(d, f32p))

    def pt_set_adjoint_mode(self, pt, mode):
        """ADJOINT_EXPECTED (derivative of the expected image, the default) or
        ADJOINT_FIXED_DECISIONS (of this realization, decisions held fixed)."""
        self._call("mfh_pt_set_adjoint_mode", self.handle, pt, mode)

    def pt_texture_gradient(self, pt, tex, shape):
        out = np.zeros(int(np.prod(shape)), np.float32)
        self._call("mfh_pt_get_texture_gradient", self.handle, pt, tex, _ptr(out, f32p))
        return out.reshape(shape)

    def pt_emission_gradient(self, pt, material_index):
        out = np.zeros(3, np.float32)
        self._call("mfh_pt_get_emission_gradient", self.handle, pt, material_index, _ptr(out, f32p))
        return out

    def pt_set_mode(self, pt, mode, use_graph=False):
        self._call("mfh_pt_set_mode", self.handle, pt, mode, 1 if use_graph else 0)

    def pt_set_tuning(self, pt, trace_block=256, stack_depth=16, queue_chunk=256, sort_bounces=0, refill_budget=0,
                      leaf_batch=1, tail_after=0):
        """Wavefront launch shape; leaf_batch (1 or 8) also applies to the megakernel;
        tail_after > 0 finishes every path in one launch after that bounce."""
        r = np.zeros(1, TUNING_DTYPE)
        r["traceBlock"], r["stackDepth"], r["queueChunk"] = trace_block, stack_depth, queue_chunk
        r["sortBounces"] = sort_bounces
        r["refillBudget"] = refill_budget
        r["leafBatch"] = leaf_batch
        r["tailAfter"] = tail_after
        self._call("mfh_pt_set_tuning", self.handle, pt, r.ctypes.data)

    def pt_get_tuning(self, pt):
        r = np.zeros(1, TUNING_DTYPE)
        self._call("mfh_pt_get_tuning", self.handle, pt, r.ctypes.data)
        return {k: int(r[k][0]) for k in TUNING_DTYPE.names}

    def pt_autotune(self, pt, iterations=2):
        r = np.zeros(1, TUNING_DTYPE)
        ms = np.zeros(1, np.float32)
        self._call("mfh_pt_autotune", self.handle, pt, iterations, r.ctypes.data, _ptr(ms, f32p))
        return {k: int(r[k][0]) for k in TUNING_DTYPE.names}, float(ms[0])

    def pt_enable_target(self, pt, target, enable=True):
        self._call("mfh_pt_enable_target", self.handle, pt, target, 1 if enable else 0)

    def pt_enable_variance(self, pt, target, enable=True):
        self._call("mfh_pt_enable_variance", self.handle, pt, target, 1 if enable else 0)

    def pt_variance(self, pt, target=TARGET_RADIANCE):
        w, h = self._pt_size[pt]
        ch = TARGET_CHANNELS[target]
        out = np.zeros(w * h * ch, np.float32)
        self._call("mfh_pt_get_variance", self.handle, pt, target, _ptr(out, f32p))
        return out.reshape(h, w, ch)

    def pt_target(self, pt, target=TARGET_RADIANCE):
        w, h = self._pt_size[pt]
        ch = TARGET_CHANNELS[target]
        out = np.zeros(w * h * ch, np.float32)
        self._call("mfh_pt_get_target", self.handle, pt, target, _ptr(out, f32p))
        return out.reshape(h, w, ch)

    def pt_iteration(self, pt):
        out = i32()
        self._call("mfh_pt_get_iteration", self.handle, pt, C.byref(out))
        return out.value

    def pt_nan_count(self, pt):
        out = C.c_uint32()
        self._call("mfh_pt_get_nan_count", self.handle, pt, C.byref(out))
        return out.value

    def pt_light_tree(self, pt):
        info = np.zeros(1, LIGHT_TREE_INFO_DTYPE)
        self._call("mfh_pt_get_light_tree", self.handle, pt, info.ctypes.data, None, None, None)
        i = info[0]
        mem = np.zeros(max(int(i["memBytes"]), 1), np.uint8)
        keys = np.zeros(max(int(i["mapCount"]), 1), np.uint64)
        codes = np.zeros(max(int(i["mapCount"]), 1), np.uint32)
        self._call("mfh_pt_get_light_tree", self.handle, pt, info.ctypes.data, mem.ctypes.data, _ptr(keys, u64p),
                   _ptr(codes, u32p))
        return i, mem[: i["memBytes"]], keys[: i["mapCount"]], codes[: i["mapCount"]]

    def timer_begin(self):
        self._call("mfh_timer_begin", self.handle)

    def timer_end(self):
        ms = C.c_float()
        self._call("mfh_timer_end", self.handle, C.byref(ms))
        return ms.value
