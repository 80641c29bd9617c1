"""A scene file (JSON + .mff, plain, DEFLATE-compressed, with packed normals)
loaded through the scene loader renders bit-identically to the same scene
built through the core API calls, on the device (triangle scenes and a room
of quads, triangles and spheres) (the built scene's render is
checked against the oracle in test_gpu_core_api)."""
import os

import numpy as np
import pytest

from mufflon_amd import core as K
from mufflon_amd import mff_writer as W
from mufflon_amd import scenes as S

pytestmark = pytest.mark.gpu

FOV_DEGREES = 40.0
BACKGROUND = ("envmap", (0.5, 0.45, 0.6))


def _scene():
    c = S.cornell_materials(12, 24, 40, 32)
    # the loader samples albedo / roughness texture files linearly and
    # environment maps nearest (json_loader.cpp:240-280, 724)
    t = c["textures"]
    c["textures"] = [(t[0][0], S.SAMPLING_LINEAR), (t[1][0], S.SAMPLING_LINEAR), (t[2][0], S.SAMPLING_NEAREST)]
    return c


def _mixed_scene():
    sc = S.mixed_scene(40, 32, instanced=False)
    return dict(sc["lods"][0], materials=sc["materials"], point_lights=sc["point_lights"], camera=sc["camera"],
                textures=[])


def _render(mi, scenario, iterations=4):
    assert mi.render_enable_renderer(0, 0), mi.core_get_dll_error()
    for name, v in (("Seed", 1), ("Max. path length", 8), ("NEE count", 1)):
        assert mi.renderer_set_parameter_int(name, v)
    assert mi.world_load_scenario(scenario), mi.core_get_dll_error()
    mi.render_enable_render_target("Radiance", False)
    for _ in range(iterations):
        ok, _ = mi.render_iterate()
        assert ok, mi.core_get_dll_error()
    return mi.mufflon_get_target_image("Radiance", False, 40, 32)


@pytest.mark.parametrize("deflate,packed,mixed", [(False, False, False), (True, False, False), (True, True, False),
                                                  (False, False, True), (True, True, True)])
def test_loaded_scene_renders_like_the_built_one(tmp_path, deflate, packed, mixed):
    print("scene file: deflate=%s packed normals=%s quads and spheres=%s" % (deflate, packed, mixed), flush=True)
    c = _mixed_scene() if mixed else _scene()
    background = None if mixed else BACKGROUND
    path = W.export_scene(c, str(tmp_path / "file"), fov_degrees=FOV_DEGREES, background=background,
                          deflate=deflate, packed_normals=packed)
    a = K.CoreInterface()
    try:
        status = a.loader_load_json(path)
        assert status == K.LOADER_SUCCESS, a.loader_get_dll_error()
        img_loaded = _render(a, a.world_find_scenario("main"))
    finally:
        a.close()
    # the same scene through the core calls; the camera angle as the loader
    # converts degrees (util/degrad.hpp), the normals as a packed file holds them
    b = K.CoreInterface()
    try:
        c2 = dict(c, normals=W.unpack_normals(W.pack_normals(c["normals"]))) if packed else c
        os.makedirs(str(tmp_path / "built"))
        vfov = np.float32(np.float32(FOV_DEGREES) * np.float32(np.pi)) / np.float32(180.0)
        scen = K.build_world(b, c2, str(tmp_path / "built"), background=background, vfov=float(vfov))
        img_built = _render(b, scen)
    finally:
        b.close()
    assert img_loaded is not None and img_loaded.shape == (32, 40, 3)
    assert np.isfinite(img_loaded).all() and img_loaded.mean() > 0
    assert np.array_equal(img_loaded.view(np.uint32), img_built.view(np.uint32))
