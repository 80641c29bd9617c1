"""ctypes binding of Mufflon's core C API subset (include/mufflon_core.h) in
the HIP library: the record layouts of core_interface.h and a thin wrapper
whose method names follow the reference's Python binding (bindings.py:322-537),
plus `build_world`, which creates a scene dict of mufflon_amd.scenes through
the API calls a scene loader makes (objects, LoD, bulk vertices, triangles
with material slots, materials, textures, camera, lights, scenario).
"""
import ctypes as C
import os

import numpy as np

from . import hip
from . import scenes as S

# ---------------------------------------------------------------- records


class Vec2(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class Vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Vec4(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class UVec3(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("z", C.c_uint32)]


class UVec4(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("z", C.c_uint32), ("w", C.c_uint32)]


class Mat3x4(C.Structure):
    _fields_ = [("v", C.c_float * 12)]


class AABB(C.Structure):
    _fields_ = [("min", Vec3), ("max", Vec3)]


class ProcessTime(C.Structure):
    _fields_ = [("cycles", C.c_uint64), ("microseconds", C.c_uint64)]


class LodMetadata(C.Structure):
    _fields_ = [("vertices", C.c_uint32), ("triangles", C.c_uint32), ("quads", C.c_uint32), ("edges", C.c_uint32),
                ("spheres", C.c_uint32)]


class LightHdl(C.Structure):
    _fields_ = [("type", C.c_uint32, 3), ("index", C.c_uint32, 29)]


class BulkLoader(C.Structure):
    _fields_ = [("type", C.c_int32), ("bytes", C.c_void_p)]  # descriptor union: FILE* or bytes


class Medium(C.Structure):
    _fields_ = [("refractionIndex", Vec2), ("absorption", Vec3)]


class LambertParams(C.Structure):
    _fields_ = [("albedo", C.c_void_p)]


class TorranceParams(C.Structure):
    _fields_ = [("roughness", C.c_void_p), ("shadowingModel", C.c_int32), ("ndf", C.c_int32), ("albedo", C.c_void_p)]


class WalterParams(C.Structure):
    _fields_ = [("roughness", C.c_void_p), ("shadowingModel", C.c_int32), ("ndf", C.c_int32),
                ("absorption", Vec3), ("refractionIndex", C.c_float)]


class EmissiveParams(C.Structure):
    _fields_ = [("radiance", C.c_void_p), ("scale", Vec3)]


class OrennayarParams(C.Structure):
    _fields_ = [("albedo", C.c_void_p), ("roughness", C.c_float)]


class MaterialParams(C.Structure):
    pass


class BlendLayer(C.Structure):
    _fields_ = [("factor", C.c_float), ("mat", C.POINTER(MaterialParams))]


class BlendParams(C.Structure):
    _fields_ = [("a", BlendLayer), ("b", BlendLayer)]


class FresnelParams(C.Structure):
    _fields_ = [("refractionIndex", Vec2), ("a", C.POINTER(MaterialParams)), ("b", C.POINTER(MaterialParams))]


class Displacement(C.Structure):
    _fields_ = [("map", C.c_void_p), ("maxMips", C.c_void_p), ("bias", C.c_float), ("scale", C.c_float)]


class MaterialUnion(C.Union):
    _fields_ = [("lambert", LambertParams), ("torrance", TorranceParams), ("walter", WalterParams),
                ("emissive", EmissiveParams), ("orennayar", OrennayarParams), ("blend", BlendParams),
                ("fresnel", FresnelParams)]


MaterialParams._fields_ = [("outerMedium", Medium), ("innerType", C.c_int32), ("alpha", C.c_void_p),
                           ("displacement", Displacement), ("inner", MaterialUnion)]

# enumerations (core_interface.h)
MATERIAL_EMISSIVE, MATERIAL_LAMBERT, MATERIAL_ORENNAYAR, MATERIAL_TORRANCE, MATERIAL_WALTER, MATERIAL_BLEND, \
    MATERIAL_FRESNEL, MATERIAL_MICROFACET = range(8)
NDF_BECKMANN, NDF_GGX, NDF_COSINE = range(3)
SHADOWING_VCAVITY, SHADOWING_SMITH = range(2)
LIGHT_POINT, LIGHT_SPOT, LIGHT_DIRECTIONAL, LIGHT_ENVMAP = range(4)
BACKGROUND_MONOCHROME, BACKGROUND_ENVMAP, BACKGROUND_SKY_HOSEK = range(3)
SAMPLING_NEAREST, SAMPLING_LINEAR = range(2)
PARAM_INT, PARAM_FLOAT, PARAM_BOOL, PARAM_ENUM = range(4)
DEVICE_CPU, DEVICE_CUDA = 1, 2
BULK_FILE, BULK_ARRAY = range(2)
NO_ANIMATION_FRAME = 0xFFFFFFFF
INVALID_MATERIAL = 0xFFFF

vp = C.c_void_p
cp = C.c_char_p
u32 = C.c_uint32
i32 = C.c_int32
Boolean = C.c_uint32

_SIGS = {
    "core_get_dll_error": (cp, []),
    "core_set_log_level": (Boolean, [i32]),
    "mufflon_initialize": (vp, []),
    "mufflon_destroy": (None, [vp]),
    "mufflon_is_cuda_available": (Boolean, []),
    "mufflon_get_target_image": (Boolean, [vp, cp, Boolean, C.POINTER(C.POINTER(C.c_float))]),
    "mufflon_get_target_image_num_channels": (Boolean, [vp, C.POINTER(C.c_int)]),
    "world_clear_all": (None, [vp]),
    "world_finalize": (Boolean, [vp, Vec3, Vec3, C.POINTER(cp)]),
    "world_reserve_objects_instances": (None, [vp, u32, u32]),
    "world_set_frame_current": (Boolean, [vp, u32]),
    "polygon_reserve": (Boolean, [vp, C.c_size_t, C.c_size_t, C.c_size_t]),
    "polygon_add_vertex": (i32, [vp, Vec3, Vec3, Vec2]),
    "polygon_add_triangle": (i32, [vp, UVec3]),
    "polygon_add_triangle_material": (i32, [vp, UVec3, C.c_uint16]),
    "polygon_add_quad": (i32, [vp, UVec4]),
    "polygon_add_quad_material": (i32, [vp, UVec4, C.c_uint16]),
    "polygon_get_quad_count": (C.c_size_t, [vp]),
    "polygon_get_face_count": (C.c_size_t, [vp]),
    "spheres_reserve": (Boolean, [vp, C.c_size_t]),
    "spheres_add_sphere": (i32, [vp, Vec3, C.c_float]),
    "spheres_add_sphere_material": (i32, [vp, Vec3, C.c_float, C.c_uint16]),
    "spheres_add_sphere_bulk": (i32, [vp, C.c_size_t, C.POINTER(BulkLoader), C.POINTER(AABB),
                                      C.POINTER(C.c_size_t)]),
    "spheres_set_material_idx": (Boolean, [vp, i32, C.c_uint16]),
    "spheres_set_material_idx_bulk": (C.c_size_t, [vp, i32, C.c_size_t, C.POINTER(BulkLoader)]),
    "spheres_get_sphere_count": (C.c_size_t, [vp]),
    "spheres_get_bounding_box": (Boolean, [vp, C.POINTER(Vec3), C.POINTER(Vec3)]),
    "polygon_add_vertex_bulk": (i32, [vp, C.c_size_t, C.POINTER(BulkLoader), C.POINTER(BulkLoader),
                                      C.POINTER(BulkLoader), C.POINTER(AABB), C.POINTER(C.c_size_t),
                                      C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "polygon_set_material_idx_bulk": (C.c_size_t, [vp, i32, C.c_size_t, C.POINTER(BulkLoader)]),
    "polygon_get_vertex_count": (C.c_size_t, [vp]),
    "polygon_get_triangle_count": (C.c_size_t, [vp]),
    "polygon_get_bounding_box": (Boolean, [vp, C.POINTER(Vec3), C.POINTER(Vec3)]),
    "world_create_object": (vp, [vp, cp, i32]),
    "world_get_object": (vp, [vp, cp]),
    "object_add_lod": (vp, [vp, u32, Boolean]),
    "world_create_instance": (vp, [vp, vp, u32]),
    "world_get_instance_count": (u32, [vp, u32]),
    "world_get_instance_by_index": (vp, [vp, u32, u32]),
    "world_get_camera_count": (C.c_size_t, [vp]),
    "world_get_camera_name": (cp, [vp]),
    "world_get_camera_position": (Boolean, [vp, C.POINTER(Vec3), u32]),
    "world_get_pinhole_camera_fov": (Boolean, [vp, C.POINTER(C.c_float)]),
    "world_find_light": (Boolean, [vp, cp, C.POINTER(LightHdl)]),
    "world_get_point_light_count": (C.c_size_t, [vp]),
    "world_get_env_light_count": (C.c_size_t, [vp]),
    "world_get_point_light_position": (Boolean, [vp, LightHdl, C.POINTER(Vec3), u32]),
    "world_get_point_light_intensity": (Boolean, [vp, LightHdl, C.POINTER(Vec3), u32]),
    "world_get_env_light_color": (Boolean, [vp, LightHdl, C.POINTER(Vec3)]),
    "world_get_env_light_scale": (Boolean, [vp, LightHdl, C.POINTER(Vec3)]),
    "world_get_scenario_count": (u32, [vp]),
    "world_get_scenario_by_index": (vp, [vp, u32]),
    "scenario_get_global_lod_level": (u32, [vp]),
    "scenario_get_resolution": (Boolean, [vp, C.POINTER(u32), C.POINTER(u32)]),
    "scenario_get_camera": (vp, [vp]),
    "scenario_is_object_masked": (Boolean, [vp, vp]),
    "scenario_get_object_lod": (u32, [vp, vp]),
    "scenario_get_point_light_count": (i32, [vp]),
    "scenario_has_envmap_light": (Boolean, [vp, vp]),
    "object_has_lod": (Boolean, [vp, u32]),
    "object_get_id": (Boolean, [vp, C.POINTER(u32)]),
    "world_get_object_name": (cp, [vp]),
    "polygon_get_face_count": (C.c_size_t, [vp]),
    "instance_set_transformation_matrix": (Boolean, [vp, vp, C.POINTER(Mat3x4), Boolean]),
    "instance_get_transformation_matrix": (Boolean, [vp, vp, C.POINTER(Mat3x4)]),
    "world_add_material": (vp, [vp, cp, C.POINTER(MaterialParams)]),
    "world_get_material_count": (i32, [vp]),
    "world_add_texture_value": (vp, [vp, C.POINTER(C.c_float), C.c_int, i32]),
    "world_add_texture": (vp, [vp, cp, i32, i32, vp, vp]),
    "world_get_texture_size": (Boolean, [vp, C.POINTER(C.c_int32 * 2)]),
    "world_add_pinhole_camera": (vp, [vp, cp, C.POINTER(Vec3), C.POINTER(Vec3), C.POINTER(Vec3), u32, C.c_float,
                                      C.c_float, C.c_float]),
    "world_get_camera": (vp, [vp, cp]),
    "world_get_camera_direction": (Boolean, [vp, C.POINTER(Vec3), u32]),
    "world_get_camera_up": (Boolean, [vp, C.POINTER(Vec3), u32]),
    "world_set_camera_position": (Boolean, [vp, Vec3, u32]),
    "world_add_light": (LightHdl, [vp, cp, i32, u32]),
    "world_add_background_light": (LightHdl, [vp, cp, i32]),
    "world_set_point_light_position": (Boolean, [vp, LightHdl, Vec3, u32]),
    "world_set_point_light_intensity": (Boolean, [vp, LightHdl, Vec3, u32]),
    "world_set_env_light_map": (Boolean, [vp, LightHdl, vp]),
    "world_set_env_light_color": (Boolean, [vp, LightHdl, Vec3]),
    "world_set_env_light_scale": (Boolean, [vp, LightHdl, Vec3]),
    "world_reserve_scenarios": (None, [vp, u32]),
    "world_create_scenario": (vp, [vp, cp]),
    "world_find_scenario": (vp, [vp, cp]),
    "world_get_current_scenario": (vp, [vp]),
    "world_finalize_scenario": (Boolean, [vp, vp, C.POINTER(cp)]),
    "world_load_scenario": (vp, [vp, vp]),
    "scene_get_bounding_box": (Boolean, [vp, C.POINTER(Vec3), C.POINTER(Vec3)]),
    "scenario_get_name": (cp, [vp]),
    "scenario_set_resolution": (Boolean, [vp, u32, u32]),
    "scenario_set_camera": (Boolean, [vp, vp, vp]),
    "scenario_add_light": (Boolean, [vp, vp, LightHdl]),
    "scenario_mask_instance": (Boolean, [vp, vp]),
    "scenario_declare_material_slot": (C.c_uint16, [vp, cp, C.c_size_t]),
    "scenario_assign_material": (Boolean, [vp, C.c_uint16, vp]),
    "scenario_get_material_slot_count": (C.c_size_t, [vp]),
    "render_get_renderer_count": (u32, [vp]),
    "render_get_renderer_variations": (u32, [vp, u32]),
    "render_get_renderer_name": (cp, [vp, u32]),
    "render_get_renderer_short_name": (cp, [vp, u32]),
    "render_get_renderer_devices": (u32, [vp, u32, u32]),
    "render_enable_renderer": (Boolean, [vp, u32, u32]),
    "render_iterate": (Boolean, [vp, C.POINTER(ProcessTime), C.POINTER(ProcessTime), C.POINTER(ProcessTime)]),
    "render_get_current_iteration": (u32, [vp]),
    "render_reset": (Boolean, [vp]),
    "render_save_screenshot": (Boolean, [vp, cp, cp, Boolean]),
    "render_get_render_target_count": (u32, [vp]),
    "render_get_render_target_name": (cp, [vp, u32]),
    "render_enable_render_target": (Boolean, [vp, cp, Boolean]),
    "render_disable_render_target": (Boolean, [vp, cp, Boolean]),
    "render_is_render_target_enabled": (Boolean, [vp, cp, Boolean]),
    "renderer_get_num_parameters": (u32, [vp]),
    "renderer_get_parameter_desc": (cp, [vp, u32, C.POINTER(i32)]),
    "renderer_set_parameter_int": (Boolean, [vp, cp, i32]),
    "renderer_get_parameter_int": (Boolean, [vp, cp, C.POINTER(i32)]),
    "renderer_set_parameter_bool": (Boolean, [vp, cp, Boolean]),
    "renderer_get_parameter_bool": (Boolean, [vp, cp, C.POINTER(Boolean)]),
    "renderer_set_parameter_float": (Boolean, [vp, cp, C.c_float]),
    "renderer_get_parameter_float": (Boolean, [vp, cp, C.POINTER(C.c_float)]),
    "world_get_texture": (vp, [vp, cp]),
    "render_set_texture_data": (Boolean, [vp, vp, C.POINTER(C.c_float)]),
    "render_set_emission_scale": (Boolean, [vp, C.c_uint16, Vec3]),
    "render_set_adjoint_mode": (Boolean, [vp, i32]),
    "render_adjoint": (Boolean, [vp, C.POINTER(C.c_float)]),
    "render_get_texture_gradient": (Boolean, [vp, vp, C.POINTER(C.c_float)]),
    "render_get_emission_gradient": (Boolean, [vp, C.c_uint16, C.POINTER(Vec3)]),
}

_core = None
_loader = None

LOADER_SUCCESS, LOADER_ERROR, LOADER_ABORT = range(3)
_LOADER_SIGS = {
    "loader_get_dll_error": (cp, []),
    "loader_set_log_level": (Boolean, [i32]),
    "loader_initialize": (vp, [vp]),
    "loader_destroy": (None, [vp]),
    "loader_load_json": (i32, [vp, cp]),
    "loader_save_scene": (i32, [vp, cp]),
    "loader_load_lod": (Boolean, [vp, vp, u32, Boolean]),
    "loader_load_object_material_indices": (Boolean, [vp, u32, C.POINTER(C.c_uint16), C.POINTER(u32)]),
    "loader_load_lod_metadatas": (Boolean, [vp, C.POINTER(LodMetadata), C.POINTER(C.c_size_t)]),
    "loader_abort": (Boolean, [vp]),
    "loader_get_loading_status": (cp, [vp]),
    "loader_profiling_enable": (None, []),
    "loader_profiling_disable": (None, []),
}


def load():
    """The HIP library with the core API signatures bound (no fallback)."""
    global _core
    if _core is None:
        lib = hip.load()
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _core = lib
    return _core


def load_loader():
    """The scene-file loader library (libmffloader.so) with its signatures
    bound; it links the HIP library, whose core API it calls."""
    global _loader
    if _loader is None:
        load()
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libmffloader.so")
        if not os.path.exists(path):
            raise RuntimeError("%s is not built: run __graft_entry__.build()" % path)
        lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
        for name, (res, args) in _LOADER_SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _loader = lib
    return _loader


def _b(s):
    return s.encode("utf-8") if isinstance(s, str) else s


def _v3(x):
    return Vec3(*(float(v) for v in x))


def bulk_array(arr):
    """A BULK_ARRAY loader over a contiguous numpy array (kept alive by the caller)."""
    return BulkLoader(BULK_ARRAY, arr.ctypes.data)


class CoreInterface:
    """One Mufflon instance (mufflon_initialize / mufflon_destroy) with the
    method names of the reference's DllInterface."""

    def __init__(self):
        self.lib = load()
        self.muffInst = self.lib.mufflon_initialize()
        if not self.muffInst:
            raise RuntimeError(self.core_get_dll_error())

    def close(self):
        if getattr(self, "_loader_inst", None):
            self.mffLoader.loader_destroy(self._loader_inst)
            self._loader_inst = None
        if self.muffInst:
            self.lib.mufflon_destroy(self.muffInst)
            self.muffInst = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def core_get_dll_error(self):
        return self.lib.core_get_dll_error().decode()

    # the scene-file loader bound to this instance (bindings.py:312-347)
    def loader(self):
        if getattr(self, "_loader_inst", None) is None:
            self.mffLoader = load_loader()
            self._loader_inst = self.mffLoader.loader_initialize(self.muffInst)
            if not self._loader_inst:
                raise RuntimeError(self.loader_get_dll_error())
        return self._loader_inst

    def loader_get_dll_error(self):
        return load_loader().loader_get_dll_error().decode()

    def loader_load_json(self, path):
        inst = self.loader()
        return self.mffLoader.loader_load_json(inst, _b(path))

    def loader_load_lod_metadatas(self, capacity=4096):
        buf = (LodMetadata * capacity)()
        n = C.c_size_t()
        if not self.mffLoader.loader_load_lod_metadatas(self.loader(), buf, C.byref(n)):
            raise RuntimeError(self.loader_get_dll_error())
        return [(m.vertices, m.triangles, m.quads, m.edges, m.spheres) for m in buf[: n.value]]

    def loader_load_object_material_indices(self, obj_id, capacity=65536):
        buf = (C.c_uint16 * capacity)()
        n = u32()
        if not self.mffLoader.loader_load_object_material_indices(self.loader(), obj_id, buf, C.byref(n)):
            raise RuntimeError(self.loader_get_dll_error())
        return list(buf[: n.value])

    def loader_get_loading_status(self):
        s = self.mffLoader.loader_get_loading_status(self.loader())
        return None if s is None else s.decode()

    # world
    def world_clear_all(self):
        self.lib.world_clear_all(self.muffInst)

    def world_finalize(self, box_min, box_max):
        msg = cp()
        ok = self.lib.world_finalize(self.muffInst, _v3(box_min), _v3(box_max), C.byref(msg))
        return bool(ok), msg.value.decode()

    def world_create_object(self, name, flags=0):
        return self.lib.world_create_object(self.muffInst, _b(name), flags)

    def world_get_object(self, name):
        return self.lib.world_get_object(self.muffInst, _b(name))

    def object_add_lod(self, obj, level=0):
        return self.lib.object_add_lod(obj, level, 0)

    def world_create_instance(self, obj, frame=NO_ANIMATION_FRAME):
        return self.lib.world_create_instance(self.muffInst, obj, frame)

    def instance_set_transformation_matrix(self, inst, m34, is_world_to_inst):
        m = Mat3x4()
        m.v[:] = [float(x) for x in np.asarray(m34, np.float32).reshape(12)]
        return bool(self.lib.instance_set_transformation_matrix(self.muffInst, inst, C.byref(m),
                                                                1 if is_world_to_inst else 0))

    def instance_get_transformation_matrix(self, inst):
        m = Mat3x4()
        if not self.lib.instance_get_transformation_matrix(self.muffInst, inst, C.byref(m)):
            raise RuntimeError(self.core_get_dll_error())
        return np.array(m.v[:], np.float32)

    def world_add_texture_value(self, values, sampling=SAMPLING_NEAREST):
        v = (C.c_float * len(values))(*[float(x) for x in values])
        return self.lib.world_add_texture_value(self.muffInst, v, len(values), sampling)

    def world_add_texture(self, path, sampling, mipmaps=0):
        return self.lib.world_add_texture(self.muffInst, _b(path), sampling, mipmaps, None, None)

    def world_add_material(self, name, params):
        return self.lib.world_add_material(self.muffInst, _b(name), C.byref(params))

    def world_add_pinhole_camera(self, name, pos, dir, up, near, far, fov_rad):
        return self.lib.world_add_pinhole_camera(self.muffInst, _b(name), C.byref(_v3(pos)), C.byref(_v3(dir)),
                                                 C.byref(_v3(up)), 1, near, far, fov_rad)

    def world_get_camera(self, name):
        return self.lib.world_get_camera(self.muffInst, _b(name))

    def world_add_light(self, name, type, count=1):
        return self.lib.world_add_light(self.muffInst, _b(name), type, count)

    def world_add_background_light(self, name, type):
        return self.lib.world_add_background_light(self.muffInst, _b(name), type)

    def world_set_point_light_position(self, light, pos, frame=0):
        return bool(self.lib.world_set_point_light_position(self.muffInst, light, _v3(pos), frame))

    def world_set_point_light_intensity(self, light, intensity, frame=0):
        return bool(self.lib.world_set_point_light_intensity(self.muffInst, light, _v3(intensity), frame))

    def world_set_env_light_map(self, light, tex):
        return bool(self.lib.world_set_env_light_map(self.muffInst, light, tex))

    def world_set_env_light_color(self, light, color):
        return bool(self.lib.world_set_env_light_color(self.muffInst, light, _v3(color)))

    def world_set_env_light_scale(self, light, scale):
        return bool(self.lib.world_set_env_light_scale(self.muffInst, light, _v3(scale)))

    def world_create_scenario(self, name):
        return self.lib.world_create_scenario(self.muffInst, _b(name))

    def world_find_scenario(self, name):
        return self.lib.world_find_scenario(self.muffInst, _b(name))

    def world_get_current_scenario(self):
        return self.lib.world_get_current_scenario(self.muffInst)

    def world_finalize_scenario(self, scenario):
        msg = cp()
        ok = self.lib.world_finalize_scenario(self.muffInst, scenario, C.byref(msg))
        return bool(ok), msg.value.decode()

    def world_load_scenario(self, scenario):
        return self.lib.world_load_scenario(self.muffInst, scenario)

    def scenario_set_resolution(self, scenario, w, h):
        return bool(self.lib.scenario_set_resolution(scenario, w, h))

    def scenario_set_camera(self, scenario, cam):
        return bool(self.lib.scenario_set_camera(self.muffInst, scenario, cam))

    def scenario_add_light(self, scenario, light):
        return bool(self.lib.scenario_add_light(self.muffInst, scenario, light))

    def scenario_declare_material_slot(self, scenario, name):
        b = _b(name)
        return self.lib.scenario_declare_material_slot(scenario, b, len(b))

    def scenario_assign_material(self, scenario, slot, mat):
        return bool(self.lib.scenario_assign_material(scenario, slot, mat))

    # rendering
    def render_get_renderer_count(self):
        return self.lib.render_get_renderer_count(self.muffInst)

    def render_get_renderer_name(self, index):
        return self.lib.render_get_renderer_name(self.muffInst, index).decode()

    def render_get_renderer_short_name(self, index):
        return self.lib.render_get_renderer_short_name(self.muffInst, index).decode()

    def render_get_renderer_variations(self, index):
        return self.lib.render_get_renderer_variations(self.muffInst, index)

    def render_get_renderer_devices(self, index, variation):
        return self.lib.render_get_renderer_devices(self.muffInst, index, variation)

    def render_enable_renderer(self, index, variation):
        return bool(self.lib.render_enable_renderer(self.muffInst, index, variation))

    def render_iterate(self):
        t = ProcessTime()
        pre = ProcessTime()
        post = ProcessTime()
        ok = self.lib.render_iterate(self.muffInst, C.byref(t), C.byref(pre), C.byref(post))
        return bool(ok), t

    def render_reset(self):
        return bool(self.lib.render_reset(self.muffInst))

    def render_get_current_iteration(self):
        return self.lib.render_get_current_iteration(self.muffInst)

    def render_save_screenshot(self, file_name, target_name, variance):
        return bool(self.lib.render_save_screenshot(self.muffInst, _b(file_name), _b(target_name),
                                                    1 if variance else 0))

    def render_get_render_target_count(self):
        return self.lib.render_get_render_target_count(self.muffInst)

    def render_get_render_target_name(self, index):
        return self.lib.render_get_render_target_name(self.muffInst, index).decode()

    def render_enable_render_target(self, name, variance=False):
        return bool(self.lib.render_enable_render_target(self.muffInst, _b(name), 1 if variance else 0))

    def render_disable_render_target(self, name, variance=False):
        return bool(self.lib.render_disable_render_target(self.muffInst, _b(name), 1 if variance else 0))

    def render_is_render_target_enabled(self, name, variance=False):
        return bool(self.lib.render_is_render_target_enabled(self.muffInst, _b(name), 1 if variance else 0))

    def renderer_set_parameter_int(self, name, value):
        return bool(self.lib.renderer_set_parameter_int(self.muffInst, _b(name), value))

    def renderer_get_parameter_int(self, name):
        v = i32()
        if not self.lib.renderer_get_parameter_int(self.muffInst, _b(name), C.byref(v)):
            raise RuntimeError(self.core_get_dll_error())
        return v.value

    def renderer_set_parameter_float(self, name, value):
        return bool(self.lib.renderer_set_parameter_float(self.muffInst, _b(name), float(value)))

    def renderer_get_parameter_float(self, name):
        v = C.c_float()
        if not self.lib.renderer_get_parameter_float(self.muffInst, _b(name), C.byref(v)):
            raise RuntimeError(self.core_get_dll_error())
        return v.value

    def renderer_set_parameter_bool(self, name, value):
        return bool(self.lib.renderer_set_parameter_bool(self.muffInst, _b(name), 1 if value else 0))

    def renderer_get_parameter_bool(self, name):
        v = Boolean()
        if not self.lib.renderer_get_parameter_bool(self.muffInst, _b(name), C.byref(v)):
            raise RuntimeError(self.core_get_dll_error())
        return bool(v.value)

    def renderer_parameters(self):
        out = []
        for i in range(self.lib.renderer_get_num_parameters(self.muffInst)):
            t = i32()
            name = self.lib.renderer_get_parameter_desc(self.muffInst, i, C.byref(t)).decode()
            out.append((name, t.value))
        return out

    # differentiable rendering (additions of this library, include/mufflon_core.h)
    def world_get_texture(self, path):
        return self.lib.world_get_texture(self.muffInst, _b(path))

    def render_set_texture_data(self, tex, texels):
        """Replace a texture's texels: (h, w, channels) in its stored format."""
        t = np.ascontiguousarray(texels, np.float32)
        if not self.lib.render_set_texture_data(self.muffInst, tex, t.ctypes.data_as(C.POINTER(C.c_float))):
            raise RuntimeError(self.core_get_dll_error())

    def render_set_emission_scale(self, slot, scale):
        if not self.lib.render_set_emission_scale(self.muffInst, int(slot), _v3(scale)):
            raise RuntimeError(self.core_get_dll_error())

    def render_set_adjoint_mode(self, mode):
        if not self.lib.render_set_adjoint_mode(self.muffInst, int(mode)):
            raise RuntimeError(self.core_get_dll_error())

    def render_adjoint(self, radiance_adjoint):
        """Replay the current Radiance image's paths with dLoss/dRadiance (h, w, 3)."""
        d = np.ascontiguousarray(radiance_adjoint, np.float32)
        if not self.lib.render_adjoint(self.muffInst, d.ctypes.data_as(C.POINTER(C.c_float))):
            raise RuntimeError(self.core_get_dll_error())

    def render_get_texture_gradient(self, tex, shape):
        """dLoss/d(texel) of a texture, shape (h, w, channels) of its stored format."""
        out = np.zeros(shape, np.float32)
        if not self.lib.render_get_texture_gradient(self.muffInst, tex, out.ctypes.data_as(C.POINTER(C.c_float))):
            raise RuntimeError(self.core_get_dll_error())
        return out

    def render_get_emission_gradient(self, slot):
        g = Vec3()
        if not self.lib.render_get_emission_gradient(self.muffInst, int(slot), C.byref(g)):
            raise RuntimeError(self.core_get_dll_error())
        return np.array([g.x, g.y, g.z], np.float32)

    def mufflon_get_target_image(self, name, variance, width, height):
        """The target as a (height, width, channels) copy, or None before a scene is loaded."""
        ptr = C.POINTER(C.c_float)()
        if not self.lib.mufflon_get_target_image(self.muffInst, _b(name), 1 if variance else 0, C.byref(ptr)):
            raise RuntimeError(self.core_get_dll_error())
        if not ptr:
            return None
        ch = C.c_int()
        self.lib.mufflon_get_target_image_num_channels(self.muffInst, C.byref(ch))
        n = width * height * ch.value
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy().reshape(height, width, ch.value)


# ---------------------------------------------------------------- scenes

_NDF = {S.NDF_GGX: NDF_GGX, S.NDF_BECKMANN: NDF_BECKMANN}
_SHADOWING = {S.SHADOWING_VCAVITY: SHADOWING_VCAVITY, S.SHADOWING_SMITH: SHADOWING_SMITH}


def write_pfm(path, texels):
    """A PFM file of (h, w) or (h, w, 1|3) float texels, little endian, rows
    in array order (the layout the core's PFM reader expects)."""
    t = np.asarray(texels, np.float32)
    if t.ndim == 3 and t.shape[2] == 1:
        t = t[:, :, 0]
    h, w = t.shape[:2]
    with open(path, "wb") as f:
        f.write(b"Pf\n" if t.ndim == 2 else b"PF\n")
        f.write(("%d %d\n-1.000000\n" % (w, h)).encode())
        f.write(np.ascontiguousarray(t[..., :3] if t.ndim == 3 else t).astype("<f4").tobytes())


def _material_params(core, m, tex_handles):
    """MaterialParams of a mufflon_amd material record; textures from
    tex_handles (id -> TextureHdl), constants as value textures."""
    p = MaterialParams()
    p.outerMedium.refractionIndex = Vec2(float(m["outerRefractionIndex"]), 0.0)
    p.outerMedium.absorption = _v3(m["outerAbsorption"])

    def albedo():
        t = int(m["albedoTexture"])
        return tex_handles[t] if t >= 0 else core.world_add_texture_value([float(x) for x in m["albedo"]])

    def roughness():
        t = int(m["roughnessTexture"])
        if t >= 0:
            return tex_handles[t]
        r = [float(x) for x in m["roughness"]]
        return core.world_add_texture_value(r[:1] if r[0] == r[1] else r)

    kind = int(m["type"])
    if kind == S.MAT_LAMBERT:
        p.innerType = MATERIAL_LAMBERT
        p.inner.lambert.albedo = albedo()
    elif kind == S.MAT_TORRANCE:
        p.innerType = MATERIAL_TORRANCE
        p.inner.torrance.albedo = albedo()
        p.inner.torrance.roughness = roughness()
        p.inner.torrance.ndf = _NDF[int(m["ndf"])]
        p.inner.torrance.shadowingModel = _SHADOWING[int(m["shadowing"])]
    elif kind in (S.MAT_WALTER, S.MAT_MICROFACET):
        p.innerType = MATERIAL_WALTER if kind == S.MAT_WALTER else MATERIAL_MICROFACET
        p.inner.walter.roughness = roughness()
        p.inner.walter.ndf = _NDF[int(m["ndf"])]
        p.inner.walter.shadowingModel = _SHADOWING[int(m["shadowing"])]
        p.inner.walter.absorption = _v3(m["absorption"])
        p.inner.walter.refractionIndex = float(m["refractionIndex"])
    elif kind == S.MAT_EMISSIVE:
        p.innerType = MATERIAL_EMISSIVE
        p.inner.emissive.radiance = core.world_add_texture_value([float(x) for x in m["radiance"]])
        p.inner.emissive.scale = _v3(m["scale"])
    else:
        raise ValueError("material type %d" % kind)
    return p


def build_world(core, scene, texture_dir, background=None, name="scene", vfov=None):
    """Create `scene` (a mufflon_amd.scenes dict with one LoD: verts, tris,
    optional quads / spheres, normals, uvs, mat_idx (per primitive),
    materials, camera, optional textures /
    point_lights / envmap) through the core API, the way a scene loader does,
    with textures written as PFM files into texture_dir. background: None,
    ("colored", color, scale) or ("envmap", scale) with the scene's envmap
    texture. vfov: the camera's vertical field of view in radians (default
    from the camera's tan_vfov). Returns the scenario handle (finalized, not
    loaded)."""
    textures = scene.get("textures") or []
    handles = {}
    for i, (texels, sampling) in enumerate(textures):
        path = os.path.join(texture_dir, "%s_tex%d.pfm" % (name, i))
        write_pfm(path, texels)
        handles[i] = core.world_add_texture(path, int(sampling))
        if not handles[i]:
            raise RuntimeError(core.core_get_dll_error())
    obj = core.world_create_object(name, 0)
    lod = core.object_add_lod(obj, 0)
    v = np.ascontiguousarray(scene["verts"], np.float32)
    n = np.ascontiguousarray(scene["normals"], np.float32)
    uv = np.ascontiguousarray(scene["uvs"], np.float32)
    pts, nrm, uvs = bulk_array(v), bulk_array(n), bulk_array(uv)
    lib = core.lib
    first = lib.polygon_add_vertex_bulk(lod, len(v), C.byref(pts), C.byref(nrm), C.byref(uvs), None, None, None,
                                        None)
    if first < 0:
        raise RuntimeError(core.core_get_dll_error())
    scen = core.world_create_scenario(name)
    mats = scene["materials"]
    slots = [core.scenario_declare_material_slot(scen, "mat%d" % i) for i in range(len(mats))]
    for i, m in enumerate(mats):
        h = core.world_add_material("mat%d" % i, _material_params(core, m, handles))
        if not h:
            raise RuntimeError(core.core_get_dll_error())
        core.scenario_assign_material(scen, slots[i], h)
    tris = np.asarray(scene["tris"], np.uint32).reshape(-1, 3)
    quads = np.asarray(scene.get("quads", np.zeros((0, 4))), np.uint32).reshape(-1, 4)
    spheres = np.ascontiguousarray(scene.get("spheres", np.zeros((0, 4))), np.float32).reshape(-1, 4)
    mat_idx = np.asarray(scene["mat_idx"])
    for t, mi in zip(tris, mat_idx):
        if lib.polygon_add_triangle_material(lod, UVec3(int(t[0]), int(t[1]), int(t[2])), int(mi)) < 0:
            raise RuntimeError(core.core_get_dll_error())
    for q, mi in zip(quads, mat_idx[len(tris):]):
        if lib.polygon_add_quad_material(lod, UVec4(*(int(x) for x in q)), int(mi)) < 0:
            raise RuntimeError(core.core_get_dll_error())
    if len(spheres):
        read = C.c_size_t()
        if lib.spheres_add_sphere_bulk(lod, len(spheres), C.byref(bulk_array(spheres)), None, C.byref(read)) < 0:
            raise RuntimeError(core.core_get_dll_error())
        sm = np.ascontiguousarray(mat_idx[len(tris) + len(quads):], np.uint16)
        lib.spheres_set_material_idx_bulk(lod, 0, len(sm), C.byref(bulk_array(sm)))
    core.world_create_instance(obj)
    cam = scene["camera"]
    if vfov is None:
        vfov = 2.0 * np.arctan(float(cam["tan_vfov"]))
    c = core.world_add_pinhole_camera("camera", cam["position"], cam["view_dir"], cam["up"], 1e-4, 1e10,
                                      float(np.float32(vfov)))
    core.scenario_set_camera(scen, c)
    core.scenario_set_resolution(scen, int(cam["width"]), int(cam["height"]))
    pl = scene.get("point_lights")
    if pl is not None:
        for i, p in enumerate(pl):
            h = core.world_add_light("point%d" % i, LIGHT_POINT, 1)
            core.world_set_point_light_position(h, p["position"])
            core.world_set_point_light_intensity(h, p["intensity"])
            core.scenario_add_light(scen, h)
    if background is not None:
        if background[0] == "colored":
            h = core.world_add_background_light("background", BACKGROUND_MONOCHROME)
            core.world_set_env_light_color(h, background[1])
            core.world_set_env_light_scale(h, background[2])
        else:
            h = core.world_add_background_light("background", BACKGROUND_ENVMAP)
            core.world_set_env_light_map(h, handles[scene["envmap"]])
            core.world_set_env_light_scale(h, background[1])
        core.scenario_add_light(scen, h)
    ok, msg = core.world_finalize_scenario(scen)
    if not ok:
        raise RuntimeError(msg)
    return scen
