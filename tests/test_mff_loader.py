"""The scene-file loader (include/mff_loader.h, libmffloader.so) on the CPU:
JSON + .mff files written by mufflon_amd.mff_writer are parsed into the core
world (objects, instances, LoDs on demand, cameras, lights, materials,
scenarios with inheritance, masks and LoD choices), and malformed files fail
with the reference's kind of message (json_loader.cpp, binary.cpp).

Without a device the final world_load_scenario of loader_load_json fails
after the LoDs it needs have been read; the tests check the world up to that
point and that the failure is that step."""
import ctypes as C
import json
import os
import struct

import numpy as np
import pytest

from mufflon_amd import core as K
from mufflon_amd import mff_writer as W
from mufflon_amd import scenes as S

CUBE_V = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], np.float32)
CUBE_T = np.array([[0, 1, 3], [0, 3, 2], [4, 6, 7], [4, 7, 5], [0, 4, 5], [0, 5, 1], [2, 3, 7], [2, 7, 6],
                   [0, 2, 6], [0, 6, 4], [1, 5, 7], [1, 7, 3]], np.uint32)


def _lod(v, t, mats):
    n = np.tile(np.array([[0, 0, 1]], np.float32), (len(v), 1))
    uv = v[:, :2].copy()
    return dict(verts=v, tris=t, normals=n, uvs=uv, mat_idx=np.asarray(mats, np.uint16))


def _translate(x, y, z):
    m = S.IDENTITY34.reshape(3, 4).copy()
    m[:, 3] = [-x, -y, -z]  # world-to-instance
    return m


def _write(tmp, deflate=False, packed=False, version="1.7", doc_edit=None, bin_version=(1, 7)):
    objects = [
        dict(name="box", lods=[_lod(CUBE_V, CUBE_T, [0] * 12), _lod(CUBE_V[:4], CUBE_T[:2], [0, 0])]),
        dict(name="tri", lods=[_lod(np.array([[0, 0, 2], [1, 0, 2], [0, 1, 2]], np.float32),
                                    np.array([[0, 1, 2]], np.uint32), [1])]),
        dict(name="lonely", lods=[_lod(CUBE_V * 0.5 + 3, CUBE_T, [1] * 12)]),
    ]
    instances = [dict(name="box_a", obj=0, world_to_inst=S.IDENTITY34),
                 dict(name="box_b", obj=0, world_to_inst=_translate(-2, 0, 0)),
                 dict(name="tri_a", obj=1, world_to_inst=S.IDENTITY34)]
    W.write_mff(os.path.join(tmp, "s.mff"), ["red", "light"], objects, instances, deflate=deflate,
                packed_normals=packed, version=bin_version)
    doc = {
        "version": version,
        "binary": "s.mff",
        "cameras": {"cam": {"type": "pinhole", "path": [[0.5, 0.5, -3]], "viewDir": [[0, 0, 1]], "fov": 40,
                            "near": 0.01, "far": 100}},
        "lights": {"pl": {"type": "point", "position": [0, 3, 0], "flux": [[12.566371, 0, 6.2831855]],
                          "scale": 2.0},
                   "sun": {"type": "spot", "position": [0, 3, 0], "direction": [0, -1, 0], "intensity": [[1, 1, 1]],
                           "width": 0.5, "falloffStart": 0.3},
                   "sky": {"type": "envmap", "scale": [0.1, 0.2, 0.3]}},
        "materials": {"redMat": {"type": "lambert", "albedo": [0.8, 0.1, 0.1]},
                      "lamp": {"type": "emissive", "radiance": [4, 4, 4], "scale": [1, 1, 1]},
                      "rough": {"type": "torrance", "albedo": [0.5, 0.5, 0.5], "roughness": [0.2, 0.4],
                                "ndf": "GGX", "shadowingModel": "smith"},
                      "veil": {"type": "orennayar", "albedo": [0.5, 0.5, 0.5]}},
        "scenarios": {
            "main": {"camera": "cam", "resolution": [64, 48], "lights": ["pl", "sky", "ghost"],
                     "materialAssignments": {"red": "redMat", "light": "lamp"},
                     "objectProperties": {"tri": {"mask": True}},
                     "instanceProperties": {"box_b": {"lod": 1}}},
            "child": {"parentScenario": "main", "resolution": [32, 24], "lod": 5,
                      "materialAssignments": {"red": "rough"}},
            "spotty": {"camera": "cam", "resolution": [8, 8], "lights": ["sun"],
                       "materialAssignments": {"red": "redMat", "light": "lamp"}},
            "veiled": {"parentScenario": "main", "materialAssignments": {"red": "veil"}},
        },
    }
    if doc_edit:
        doc_edit(doc)
    path = os.path.join(tmp, "s.json")
    with open(path, "w") as f:
        json.dump(doc, f)
    return path


def _load(path):
    mi = K.CoreInterface()
    status = mi.loader_load_json(path)
    err = mi.loader_get_dll_error()
    return mi, status, err


def _loaded_up_to_device(status, err):
    """Either loaded (a device is present) or failed only at the device."""
    assert status == K.LOADER_SUCCESS or (status == K.LOADER_ERROR and "Cannot load the default scenario 'main'"
                                          in err), err


@pytest.mark.parametrize("deflate,packed", [(False, False), (True, False), (True, True), (False, True)])
def test_world_from_scene_file(tmp_path, deflate, packed):
    mi, status, err = _load(_write(str(tmp_path), deflate, packed))
    _loaded_up_to_device(status, err)
    lib = mi.lib
    box, tri, lonely = (mi.world_get_object(n) for n in ("box", "tri", "lonely"))
    assert box and tri and lonely
    ids = []
    for o in (box, tri, lonely):
        i = C.c_uint32()
        assert lib.object_get_id(o, C.byref(i))
        ids.append(i.value)
    assert ids == [0, 1, 2]
    # three file instances + the default instance of the object without one
    assert lib.world_get_instance_count(mi.muffInst, K.NO_ANIMATION_FRAME) == 4
    m = mi.instance_get_transformation_matrix(lib.world_get_instance_by_index(mi.muffInst, 1, K.NO_ANIMATION_FRAME))
    assert np.allclose(m.reshape(3, 4)[:, 3], [-2, 0, 0])
    # LoDs loaded on demand for the default scenario: box_a LoD 0, box_b LoD 1
    # (instance property), tri masked (not loaded), lonely LoD 0
    assert [bool(lib.object_has_lod(box, 0)), bool(lib.object_has_lod(box, 1)), bool(lib.object_has_lod(tri, 0)),
            bool(lib.object_has_lod(lonely, 0))] == [True, True, False, True]
    lod0 = lib.object_add_lod(box, 0, 0)
    lod1 = lib.object_add_lod(box, 1, 0)
    assert (lib.polygon_get_vertex_count(lod0), lib.polygon_get_triangle_count(lod0)) == (8, 12)
    assert (lib.polygon_get_vertex_count(lod1), lib.polygon_get_triangle_count(lod1)) == (4, 2)
    lo, hi = K.Vec3(), K.Vec3()
    assert lib.polygon_get_bounding_box(lib.object_add_lod(lonely, 0, 0), C.byref(lo), C.byref(hi))
    assert np.allclose([lo.x, lo.y, lo.z, hi.x, hi.y, hi.z], [3, 3, 3, 3.5, 3.5, 3.5])
    # LoD sizes and unique material indices read from the file
    assert mi.loader_load_lod_metadatas() == [(8, 12, 0, 0, 0), (4, 2, 0, 0, 0), (3, 1, 0, 0, 0), (8, 12, 0, 0, 0)]
    assert mi.loader_load_object_material_indices(0) == [0]
    assert mi.loader_load_object_material_indices(2) == [1]
    mi.close()


def test_scenarios_cameras_lights_materials(tmp_path):
    mi, status, err = _load(_write(str(tmp_path)))
    _loaded_up_to_device(status, err)
    lib = mi.lib
    # "spotty" uses a spot light and "veiled" an Oren-Nayar material: not created
    names = [lib.scenario_get_name(lib.world_get_scenario_by_index(mi.muffInst, i)).decode()
             for i in range(lib.world_get_scenario_count(mi.muffInst))]
    assert names == ["main", "child"]
    assert not mi.world_find_scenario("spotty") and not mi.world_find_scenario("veiled")
    main, child = mi.world_find_scenario("main"), mi.world_find_scenario("child")
    w, h = C.c_uint32(), C.c_uint32()
    assert lib.scenario_get_resolution(main, C.byref(w), C.byref(h)) and (w.value, h.value) == (64, 48)
    assert lib.scenario_get_resolution(child, C.byref(w), C.byref(h)) and (w.value, h.value) == (32, 24)
    assert lib.scenario_get_global_lod_level(main) == 0 and lib.scenario_get_global_lod_level(child) == 5
    tri = mi.world_get_object("tri")
    assert lib.scenario_is_object_masked(main, tri) and lib.scenario_is_object_masked(child, tri)  # inherited
    # point light and background of main ("ghost" is unknown and skipped)
    assert lib.scenario_get_point_light_count(main) == 1
    assert not lib.scenario_has_envmap_light(mi.muffInst, main)  # "sky" is monochrome (no map)
    assert lib.scenario_get_material_slot_count(main) == 2
    # pinhole camera: fov degrees -> radians in float arithmetic
    cam = lib.scenario_get_camera(main)
    fov = C.c_float()
    assert lib.world_get_pinhole_camera_fov(cam, C.byref(fov))
    assert np.float32(fov.value) == np.float32(np.float32(40.0) * np.float32(np.pi)) / np.float32(180.0)
    # flux [W] over 4 pi sr times the scale
    pl = K.LightHdl()
    assert lib.world_find_light(mi.muffInst, b"pl", C.byref(pl))
    iv = K.Vec3()
    assert lib.world_get_point_light_intensity(mi.muffInst, pl, C.byref(iv), 0)
    four_pi = np.float32(4.0) * np.float32(np.pi)
    assert np.allclose([iv.x, iv.y, iv.z], [np.float32(12.566371) / four_pi * 2, 0, np.float32(6.2831855) / four_pi * 2],
                       rtol=1e-6)
    sky = K.LightHdl()
    assert lib.world_find_light(mi.muffInst, b"sky", C.byref(sky))
    col = K.Vec3()
    assert lib.world_get_env_light_color(mi.muffInst, sky, C.byref(col))
    assert np.allclose([col.x, col.y, col.z], [0.1, 0.2, 0.3])
    assert lib.world_get_point_light_count(mi.muffInst) == 1  # the spot light is not added
    assert lib.world_get_material_count(mi.muffInst) == 3  # lambert, emissive, torrance
    mi.close()


def test_lod_clamped_to_last_and_default_scenario_first(tmp_path):
    def edit(doc):
        doc["scenarios"] = {"deep": dict(doc["scenarios"]["main"], lod=7, instanceProperties={}),
                            "main": doc["scenarios"]["main"]}
    mi, status, err = _load(_write(str(tmp_path), doc_edit=edit))
    # no defaultScenario key: the first scenario is loaded; LoD 7 of a
    # two-LoD object is its last one
    assert status == K.LOADER_SUCCESS or "Cannot load the default scenario 'deep'" in err, err
    box = mi.world_get_object("box")
    assert mi.lib.object_has_lod(box, 1) and not mi.lib.object_has_lod(box, 0)
    mi.close()


def test_version_before_unique_material_indices(tmp_path):
    mi, status, err = _load(_write(str(tmp_path), version="1.6", bin_version=(1, 6)))
    _loaded_up_to_device(status, err)
    assert mi.loader_load_object_material_indices(0) == []
    mi.close()


def test_world_to_instance_before_version_1_4(tmp_path):
    # files before 1.4 store instance-to-world matrices
    mi, status, err = _load(_write(str(tmp_path), version="1.3"))
    assert status == K.LOADER_SUCCESS or "Cannot load the default scenario" in err, err
    m = mi.instance_get_transformation_matrix(mi.lib.world_get_instance_by_index(mi.muffInst, 1,
                                                                                  K.NO_ANIMATION_FRAME))
    assert np.allclose(m.reshape(3, 4)[:, 3], [2, 0, 0])
    mi.close()


@pytest.mark.parametrize("edit,message", [
    (lambda d: d["scenarios"]["main"].pop("camera"), 'is missing key "camera"'),
    (lambda d: d["scenarios"]["main"].update(camera=3), 'has invalid type'),
    (lambda d: d["scenarios"]["main"].update(camera="nope"), "Camera 'nope' does not exist"),
    (lambda d: d["scenarios"]["main"]["materialAssignments"].update(red="nope"), "Unknown material name 'nope'"),
    (lambda d: d["scenarios"]["main"]["materialAssignments"].pop("light"), 'is missing key "light"'),
    (lambda d: d.update(binary="missing.mff"), "does not exist"),
    (lambda d: d["scenarios"]["main"].update(lights=["sun"]), "The default scenario 'main' uses the spot light"),
    (lambda d: d["lights"]["pl"].pop("flux") and d["lights"]["pl"].update(temperature=[3000]),
     "black-body temperatures are not supported"),
    (lambda d: d["materials"]["rough"].update(ndf="Weird"), "Unknown normal distribution function 'Weird'"),
    (lambda d: d["scenarios"]["child"].update(parentScenario="child"), "Loop in scenario inheritance"),
    (lambda d: d.pop("materials"), 'is missing key "materials"'),
])
def test_errors(tmp_path, edit, message):
    mi, status, err = _load(_write(str(tmp_path), doc_edit=edit))
    assert status == K.LOADER_ERROR
    assert message in err, err
    mi.close()


def test_json_syntax_error_names_offset_and_line(tmp_path):
    p = _write(str(tmp_path))
    text = open(p).read()
    with open(p, "w") as f:
        f.write(text[:40] + "\n,\n" + text[40:])
    mi, status, err = _load(p)
    assert status == K.LOADER_ERROR and "Parser error" in err and "(line 1)" in err, err
    mi.close()


def test_corrupt_binary(tmp_path):
    p = _write(str(tmp_path))
    b = bytearray(open(os.path.join(str(tmp_path), "s.mff"), "rb").read())
    b[0:4] = b"Matz"
    open(os.path.join(str(tmp_path), "s.mff"), "wb").write(bytes(b))
    mi, status, err = _load(p)
    assert status == K.LOADER_ERROR and "Invalid materials header magic constant" in err, err
    mi.close()


@pytest.mark.parametrize("deflate", [False, True])
def test_quads_and_spheres_load(tmp_path, deflate):
    # a LoD of triangles, quads and spheres: quads follow the triangles in the
    # face order with their material indices after the triangles' ones
    # (binary.cpp:711-733); spheres come with their own material stream, one
    # deflated block with the centres and radii when compressed (binary.cpp:539-567)
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1]], np.float32)
    t = np.array([[0, 1, 4], [1, 5, 4]], np.uint32)
    q = np.array([[0, 1, 2, 3], [0, 3, 4, 5], [1, 2, 5, 4]], np.uint32)
    sp = np.array([[3, 0, 0, 0.5], [-2, 1, 0, 1.0], [0, 4, 0, 0.25], [0, 0, -3, 2.0]], np.float32)
    lod = _lod(v, t, [0, 1, 1, 0, 1])
    lod.update(quads=q, spheres=sp, sphere_mat_idx=np.array([1, 0, 1, 1], np.uint16))
    objects = [dict(name="mixed", lods=[lod]), dict(name="balls", lods=[dict(verts=np.zeros((0, 3), np.float32),
                                                                            spheres=sp[:2], sphere_mat_idx=[0, 0])])]
    W.write_mff(os.path.join(str(tmp_path), "s.mff"), ["red", "light"], objects,
                [dict(name="m", obj=0, world_to_inst=S.IDENTITY34), dict(name="b", obj=1, world_to_inst=S.IDENTITY34)],
                deflate=deflate)
    doc = {"version": "1.7", "binary": "s.mff",
           "cameras": {"cam": {"type": "pinhole", "path": [[0, 0, -6]], "viewDir": [[0, 0, 1]], "fov": 40}},
           "lights": {"pl": {"type": "point", "position": [0, 3, 0], "flux": [[1, 1, 1]]}},
           "materials": {"redMat": {"type": "lambert", "albedo": [0.8, 0.1, 0.1]},
                         "lamp": {"type": "emissive", "radiance": [4, 4, 4], "scale": [1, 1, 1]}},
           "scenarios": {"main": {"camera": "cam", "resolution": [16, 16], "lights": ["pl"],
                                  "materialAssignments": {"red": "redMat", "light": "lamp"}}}}
    path = os.path.join(str(tmp_path), "s.json")
    with open(path, "w") as f:
        json.dump(doc, f)
    mi, status, err = _load(path)
    _loaded_up_to_device(status, err)
    lib = mi.lib
    obj = mi.world_get_object("mixed")
    h = lib.object_add_lod(obj, 0, 0)
    assert lib.polygon_get_triangle_count(h) == 2 and lib.polygon_get_quad_count(h) == 3
    assert lib.polygon_get_face_count(h) == 5 and lib.spheres_get_sphere_count(h) == 4
    lo, hi = K.Vec3(), K.Vec3()
    assert lib.spheres_get_bounding_box(h, C.byref(lo), C.byref(hi))
    # the sphere box grows by the object's box (sphere.cpp:87-115)
    assert (lo.x, lo.y, lo.z, hi.x, hi.y, hi.z) == (-3.0, -2.0, -5.0, 3.5, 4.25, 1.0)
    assert mi.loader_load_object_material_indices(0) == [0, 1]
    assert lib.spheres_get_sphere_count(lib.object_add_lod(mi.world_get_object("balls"), 0, 0)) == 2
    mi.close()


@pytest.mark.parametrize("deflate,packed", [(False, False), (True, True)])
def test_exported_scene_loads(tmp_path, deflate, packed):
    # a scene of mufflon_amd.scenes written with export_scene reads back with
    # its triangles, materials, lights and camera
    c = S.cornell_materials(12, 24, 40, 32)
    path = W.export_scene(c, str(tmp_path), fov_degrees=40.0, background=("envmap", (0.5, 0.45, 0.6)),
                          deflate=deflate, packed_normals=packed)
    mi, status, err = _load(path)
    _loaded_up_to_device(status, err)
    lib = mi.lib
    obj = mi.world_get_object("scene")
    lod = lib.object_add_lod(obj, 0, 0)
    assert lib.polygon_get_triangle_count(lod) == len(c["tris"])
    assert lib.polygon_get_vertex_count(lod) == len(c["verts"])
    assert lib.world_get_material_count(mi.muffInst) == len(c["materials"])
    assert lib.world_get_point_light_count(mi.muffInst) == 2
    assert lib.world_get_env_light_count(mi.muffInst) == 1
    main = mi.world_find_scenario("main")
    assert lib.scenario_has_envmap_light(mi.muffInst, main)
    assert lib.scenario_get_material_slot_count(main) == len(c["materials"])
    assert mi.loader_load_object_material_indices(0) == sorted(set(int(m) for m in c["mat_idx"]))
    mi.close()


def test_normal_packing_closed_forms():
    # axes and diagonals pack to the documented u, v; unpacking inverts packing
    # of packed values exactly and approximates the input within the grid
    axes = np.array([[0, 0, 1], [0, 0, -1], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0]], np.float32)
    p = W.pack_normals(axes)
    assert p[0] == 0  # u = v = 0
    assert p[2] == 32767 and p[4] == (32767 << 16)
    assert p[1] == (32767 | (32767 << 16))  # lower hemisphere pole: u = v = 1 (signs of +0)
    rng = np.random.default_rng(3)
    n = rng.normal(size=(2000, 3)).astype(np.float32)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    u = W.unpack_normals(W.pack_normals(n))
    assert np.abs(u - n).max() < 2e-4
    assert np.array_equal(W.pack_normals(u), W.pack_normals(n))
    assert np.allclose(np.linalg.norm(u, axis=1), 1.0, atol=1e-6)


def test_loader_status_and_abort_without_load():
    mi = K.CoreInterface()
    mi.loader()
    assert mi.loader_get_loading_status() == ""
    assert not mi.mffLoader.loader_abort(mi.loader())
    assert "No loader running" in mi.loader_get_dll_error()
    assert mi.mffLoader.loader_save_scene(mi.loader(), b"/tmp/x.json") == K.LOADER_ERROR
    mi.close()
