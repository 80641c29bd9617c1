"""The core C API (include/mufflon_core.h) end to end on the device: a scene
built through the loader calls, rendered with render_iterate and read with
mufflon_get_target_image, against the oracle on the same inputs; both
renderer variations; variance targets; PFM screenshots; quads and spheres."""
import ctypes as C
import os

import numpy as np
import pytest

import pt_util as P
import scene_util as U
from mufflon_amd import core as K
from mufflon_amd import hip
from mufflon_amd import scenes as S

pytestmark = pytest.mark.gpu

RTOL = 1e-4
ATOL = 1e-6


def _tanf(x):
    libm = C.CDLL("libm.so.6")
    libm.tanf.restype = C.c_float
    libm.tanf.argtypes = [C.c_float]
    return np.float32(libm.tanf(np.float32(x)))


def _oracle_scene_camera(mi, c):
    """The camera as the core stores it (set_view_dir, and tanVFov =
    tanf(vFov / 2) of the float vFov it was given)."""
    cam = mi.world_get_camera("camera")
    d, u = K.Vec3(), K.Vec3()
    mi.lib.world_get_camera_direction(cam, C.byref(d), 0)
    mi.lib.world_get_camera_up(cam, C.byref(u), 0)
    vfov = np.float32(2.0 * np.arctan(float(c["camera"]["tan_vfov"])))
    return dict(c["camera"], view_dir=np.array([d.x, d.y, d.z], np.float32),
                up=np.array([u.x, u.y, u.z], np.float32), tan_vfov=_tanf(vfov / np.float32(2)))


def _oracle_scene(oracle, mi, c, bg):
    """Oracle scene with the camera as the core stores it."""
    return P.oracle_pt_scene_materials(oracle, dict(c, camera=_oracle_scene_camera(mi, c)), bg)


def _compare(g, o, what):
    bad = np.abs(g - o) > RTOL * np.abs(o) + ATOL
    print("%s: bit-exact=%s, out of tolerance=%d" % (what, np.array_equal(g.view(np.uint32), o.view(np.uint32)),
                                                      int(bad.sum())), flush=True)
    assert not bad.any(), what


def test_core_api_render_matches_oracle(oracle, tmp_path):
    print("core api: material zoo through the loader calls", flush=True)
    c = S.cornell_materials(12, 24, 40, 32)
    mi = K.CoreInterface()
    try:
        scen = K.build_world(mi, c, str(tmp_path), background=("envmap", (0.5, 0.45, 0.6)))
        assert mi.render_enable_renderer(0, 0)
        for name, v in (("Seed", 1), ("Max. path length", 8), ("NEE count", 1)):
            assert mi.renderer_set_parameter_int(name, v)
        assert mi.world_load_scenario(scen), mi.core_get_dll_error()
        for t in ("Radiance", "Hit ID", "Depth"):
            mi.render_enable_render_target(t, t == "Radiance")
        for _ in range(4):
            ok, t = mi.render_iterate()
            assert ok, mi.core_get_dll_error()
        assert mi.render_get_current_iteration() == 4
        W, H = 40, 32
        rad = mi.mufflon_get_target_image("Radiance", False, W, H)
        var = mi.mufflon_get_target_image("Radiance", True, W, H)
        depth = mi.mufflon_get_target_image("Depth", False, W, H)
        with pytest.raises(RuntimeError):
            mi.mufflon_get_target_image("Depth", True, W, H)
        ps = _oracle_scene(oracle, mi, c, P.background_of(c, "envmap"))
        mean, ovar, _ = oracle.pt_render_targets(ps, 1, 0, 8, 1, 4, [0, 1, 3], variance=[0])
        _compare(rad, mean[0], "core radiance")
        _compare(var, ovar[0], "core radiance variance")
        _compare(depth, mean[3], "core depth")
        # the wavefront variation renders the same image (a renderer switch
        # starts a fresh output: targets are enabled again)
        assert mi.render_enable_renderer(0, 1)
        for name, v in (("Seed", 1), ("Max. path length", 8), ("NEE count", 1)):
            assert mi.renderer_set_parameter_int(name, v)
        mi.render_enable_render_target("Radiance", False)
        for _ in range(4):
            assert mi.render_iterate()[0], mi.core_get_dll_error()
        wf = mi.mufflon_get_target_image("Radiance", False, W, H)
        assert np.array_equal(wf, rad)
        # a screenshot with tag replacement (interface.cpp:168-195)
        assert mi.render_save_screenshot(str(tmp_path / "#scenario-#shortrenderer-#iteration-#target"), "Radiance",
                                         False), mi.core_get_dll_error()
        path = tmp_path / "scene-PT-4-Radiance.pfm"
        raw = path.read_bytes()
        head = b"PF\n40 32\n-1.000000\n"
        assert raw.startswith(head)
        img = np.frombuffer(raw[len(head):], "<f4").reshape(H, W, 3)
        assert np.array_equal(img, wf)
        # a parameter change restarts the render (on_renderer_parameter_changed)
        assert mi.renderer_set_parameter_int("Seed", 2)
        assert mi.render_get_current_iteration() == 0
        assert mi.render_iterate()[0]
        assert mi.render_get_current_iteration() == 1
    finally:
        mi.close()


def test_core_api_wireframe_renderer(oracle, tmp_path):
    """The "Wireframe"/"WF" renderer through the core API: parameters, the
    Border target with variance, against the oracle's wireframe on the
    camera the core stores."""
    print("core api: wireframe renderer", flush=True)
    c = S.cornell_materials(12, 24, 40, 32)
    mi = K.CoreInterface()
    try:
        scen = K.build_world(mi, c, str(tmp_path), background=("colored", (0.3, 0.4, 0.6), (0.8, 0.8, 0.8)))
        assert mi.render_enable_renderer(1, 0)
        assert mi.renderer_set_parameter_float("Line width", 1.5)
        assert mi.renderer_set_parameter_int("Maximum trace depth", 50)
        assert mi.world_load_scenario(scen), mi.core_get_dll_error()
        mi.render_enable_render_target("Border", True)
        for _ in range(3):
            ok, _ = mi.render_iterate()
            assert ok, mi.core_get_dll_error()
        W, H = 40, 32
        border = mi.mufflon_get_target_image("Border", False, W, H)
        var = mi.mufflon_get_target_image("Border", True, W, H)
        ps = _oracle_scene(oracle, mi, c, P.background_of(c, "colored"))
        o, ov, _ = oracle.wireframe_render(ps, 3, line_width=1.5, max_trace_depth=50, variance=True)
        _compare(border, o, "core border")
        _compare(var, ov, "core border variance")
        assert np.array_equal(border.view(np.uint32), o.view(np.uint32))
        assert mi.render_save_screenshot(str(tmp_path / "#shortrenderer-#target"), "Border", False)
        assert (tmp_path / "WF-Border.pfm").exists()
    finally:
        mi.close()


def test_core_api_mixed_scene_matches_oracle(oracle, tmp_path):
    """Quads and spheres through the loader calls (polygon_add_quad_material,
    spheres_add_sphere_bulk, spheres_set_material_idx_bulk): the room of
    bilinear quads, triangle boxes and spheres with a quad light and a sphere
    light renders as the oracle renders it."""
    print("core api: quads and spheres", flush=True)
    sc = S.mixed_scene(40, 32, instanced=False)
    lod = sc["lods"][0]
    c = dict(lod, materials=sc["materials"], point_lights=sc["point_lights"], camera=sc["camera"], textures=[])
    mi = K.CoreInterface()
    try:
        scen = K.build_world(mi, c, str(tmp_path))
        assert mi.render_enable_renderer(0, 0)
        for name, v in (("Seed", 1), ("Max. path length", 8), ("NEE count", 1)):
            assert mi.renderer_set_parameter_int(name, v)
        assert mi.world_load_scenario(scen), mi.core_get_dll_error()
        for t in ("Radiance", "Depth"):
            mi.render_enable_render_target(t, False)
        for _ in range(3):
            assert mi.render_iterate()[0], mi.core_get_dll_error()
        W, H = 40, 32
        rad = mi.mufflon_get_target_image("Radiance", False, W, H)
        depth = mi.mufflon_get_target_image("Depth", False, W, H)
        cam = _oracle_scene_camera(mi, c)
        _, lods = U.oracle_scene(oracle, [(lod["verts"], lod["tris"], lod["quads"], lod["spheres"])],
                                 sc["inst_lod"], sc["world_to_inst"])
        _, box = oracle.instance_boxes(sc["world_to_inst"], np.stack([L["box"] for L in lods]), sc["inst_lod"])
        box = np.asarray(box, np.float32)
        lt = hip.light_tree_build(S.emissive_area_lights_ex(lod, sc["materials"]), sc["point_lights"], box,
                                  sc["materials"])
        ps = P.oracle_pt_scene_general(oracle, dict(sc, camera=cam), lt, box)
        mean, _ = oracle.pt_render_targets(ps, 1, 0, 8, 1, 3, [0, 3])
        _compare(rad, mean[0], "core mixed radiance")
        _compare(depth, mean[3], "core mixed depth")
        assert mean[0].mean() > 0.01
    finally:
        mi.close()


def test_core_api_lighttracer(oracle, tmp_path):
    """The "Lighttracer"/"LT" renderer through the core API: its parameters,
    the Radiance (with variance) and Density targets, against the oracle's
    light tracer on the camera the core stores (float atomics: 1e-4 relative
    plus 1e-6 of the largest value)."""
    print("core api: light tracer", flush=True)
    c = P.cornell(12, 24, 40)
    mi = K.CoreInterface()
    try:
        scen = K.build_world(mi, c, str(tmp_path))
        assert mi.render_enable_renderer(2, 0)
        for name, v in (("Seed", 2), ("Max. path length", 6)):
            assert mi.renderer_set_parameter_int(name, v)
        assert mi.world_load_scenario(scen), mi.core_get_dll_error()
        mi.render_enable_render_target("Radiance", True)
        mi.render_enable_render_target("Density", False)
        for _ in range(3):
            assert mi.render_iterate()[0], mi.core_get_dll_error()
        W, H = 40, 40
        got = {"radiance": mi.mufflon_get_target_image("Radiance", False, W, H),
               "radiance variance": mi.mufflon_get_target_image("Radiance", True, W, H),
               "density": mi.mufflon_get_target_image("Density", False, W, H)}
        ps = P.oracle_pt_scene(oracle, dict(c, camera=_oracle_scene_camera(mi, c)))
        ref, rvar = oracle.lt_render(ps, 2, 0, 6, 3, density=True, variance=True)
        want = {"radiance": ref[..., :3], "radiance variance": rvar[..., :3], "density": ref[..., 4:5]}
        for k in got:
            o = want[k]
            g = got[k].reshape(o.shape)
            tol = 1e-4 * np.abs(o) + 1e-6 * max(float(np.abs(o).max()), 1e-30)
            print("%s: out of tolerance %d" % (k, int((np.abs(g - o) > tol).sum())), flush=True)
            assert (np.abs(g - o) <= tol).all(), k
        assert ref[..., :3].mean() > 0
        assert mi.render_save_screenshot(str(tmp_path / "#shortrenderer-#target"), "Radiance", False)
        assert (tmp_path / "LT-Radiance.pfm").exists()
    finally:
        mi.close()
