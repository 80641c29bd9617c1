"""Writer of Mufflon's scene-file format (a JSON description plus a binary
.mff file; doc/FileFormat.md of the reference), so scenes made with
mufflon_amd.scenes can be loaded through the scene loader
(libmffloader.so, include/mff_loader.h) or by Mufflon itself.

The binary layout follows the format description section by section: the
materials header, the objects section (flags, jump table, per object the
name, flags, keyframes, box, LoD jump table and - from version 1.7 - the
sorted unique material indices), LoDs (counts, vertex data, triangles,
quads, face material indices, spheres with their material indices; DEFLATE
streams when compressed) and the instances with world-to-instance matrices.
"""
import json
import os
import struct
import zlib

import numpy as np

from . import scenes as S
from .core import write_pfm

FLAG_DEFLATE = 1
FLAG_COMPRESSED_NORMALS = 2
NO_FRAME = 0xFFFFFFFF


def pack_normals(normals):
    """Octahedral 32-bit normal packing of the format description ("Normal
    Compression"): u, v in [-1, 1] to int16 in the low / high half."""
    n = np.asarray(normals, np.float64)
    l1 = np.abs(n).sum(axis=1)
    l1[l1 == 0] = 1.0
    u = n[:, 0] / l1
    v = n[:, 1] / l1
    lower = n[:, 2] < 0
    su = np.where(n[:, 0] >= 0, 1.0, -1.0)
    sv = np.where(n[:, 1] >= 0, 1.0, -1.0)
    u2 = (1.0 - np.abs(n[:, 1]) / l1) * su
    v2 = (1.0 - np.abs(n[:, 0]) / l1) * sv
    u = np.where(lower, u2, u)
    v = np.where(lower, v2, v)
    qu = np.floor(u * 32767.0 + 0.5).astype(np.int64) & 0xFFFF
    qv = np.floor(v * 32767.0 + 0.5).astype(np.int64) & 0xFFFF
    return (qu | (qv << 16)).astype(np.uint32)


def unpack_normals(packed):
    """The loader's decoding of packed normals in float32 arithmetic (the
    normals a loaded LoD holds)."""
    p = np.asarray(packed, np.uint32)
    u = (p & 0xFFFF).astype(np.uint16).view(np.int16).astype(np.float32) / np.float32(32767.0)
    v = (p >> 16).astype(np.uint16).view(np.int16).astype(np.float32) / np.float32(32767.0)
    z = np.float32(1.0) - np.abs(u) - np.abs(v)
    su = np.where(u >= 0, np.float32(1.0), np.float32(-1.0)).astype(np.float32)
    sv = np.where(v >= 0, np.float32(1.0), np.float32(-1.0)).astype(np.float32)
    x = np.where(z >= 0, u, (np.float32(1.0) - np.abs(v)) * su).astype(np.float32)
    y = np.where(z >= 0, v, (np.float32(1.0) - np.abs(u)) * sv).astype(np.float32)
    ln = np.sqrt(x * x + y * y + z * z).astype(np.float32)
    return np.stack([x / ln, y / ln, z / ln], axis=1).astype(np.float32)


def _u32(v):
    return struct.pack("<I", v)


def _u64(v):
    return struct.pack("<Q", v)


def _str(s):
    b = s.encode("utf-8")
    return _u32(len(b)) + b


def _block(data, deflate):
    if not deflate:
        return data
    packed = zlib.compress(data)
    return _u32(len(packed)) + _u32(len(data)) + packed


def _lod_bytes(lod, deflate, packed_normals):
    v = np.ascontiguousarray(lod["verts"], "<f4").reshape(-1, 3)
    t = np.ascontiguousarray(lod.get("tris", np.zeros((0, 3))), "<u4").reshape(-1, 3)
    q = np.ascontiguousarray(lod.get("quads", np.zeros((0, 4))), "<u4").reshape(-1, 4)
    sp = np.ascontiguousarray(lod.get("spheres", np.zeros((0, 4))), "<f4").reshape(-1, 4)
    n = lod.get("normals")
    n = np.zeros_like(v) if n is None else np.ascontiguousarray(n, "<f4")
    uv = lod.get("uvs")
    uv = np.zeros((len(v), 2), "<f4") if uv is None else np.ascontiguousarray(uv, "<f4")
    mats = np.ascontiguousarray(lod.get("mat_idx", np.zeros(len(t) + len(q))), "<u2")
    sphere_mats = np.ascontiguousarray(lod.get("sphere_mat_idx", np.zeros(len(sp))), "<u2")
    assert len(mats) == len(t) + len(q) and len(sphere_mats) == len(sp)
    out = bytearray(b"LOD_")
    out += struct.pack("<8I", len(t), len(q), len(sp), len(v), 0, 0, 0, 0)
    if len(v):
        normal_bytes = pack_normals(n).astype("<u4").tobytes() if packed_normals else n.tobytes()
        out += _block(v.tobytes() + normal_bytes + uv.tobytes(), deflate)
    if len(t):
        out += _block(t.tobytes(), deflate)
    if len(q):
        out += _block(q.tobytes(), deflate)
    if len(t) + len(q):
        out += _block(mats.tobytes(), deflate)
    if len(sp):
        # centre and radius, then the material indices: one stream when deflated
        if deflate:
            out += _block(sp.tobytes() + sphere_mats.tobytes(), True)
        else:
            out += sp.tobytes() + sphere_mats.tobytes()
    return bytes(out)


def write_mff(path, material_names, objects, instances, deflate=False, packed_normals=False, version=(1, 7)):
    """Write the binary file. objects: dicts with name, lods (list of dicts
    with verts, tris, normals, uvs, mat_idx - one per triangle, then one per
    quad - and optional quads (n x 4 vertex indices), spheres (n x 4: centre,
    radius) and sphere_mat_idx), optional box (6 floats; default
    the union of the LoDs' vertex bounds), flags, keyframe. instances: dicts
    with name, obj (index), world_to_inst (3x4), optional keyframe."""
    flags = (FLAG_DEFLATE if deflate else 0) | (FLAG_COMPRESSED_NORMALS if packed_normals else 0)
    mats = b"".join(_str(m) for m in material_names)
    head_len = 4 + 8 + 4 + len(mats)
    out = bytearray(b"Mats" + _u64(head_len) + _u32(len(material_names)) + mats)
    # objects section: header + object jump table, then the objects
    objs_start = len(out)
    obj_blobs = []
    for o in objects:
        lods = [_lod_bytes(L, deflate, packed_normals) for L in o["lods"]]
        box = o.get("box")
        if box is None:
            lo, hi = [], []
            for L in o["lods"]:
                v = np.asarray(L["verts"], np.float32).reshape(-1, 3)
                sp = np.asarray(L.get("spheres", np.zeros((0, 4))), np.float32).reshape(-1, 4)
                lo += [v, sp[:, :3] - sp[:, 3:]]
                hi += [v, sp[:, :3] + sp[:, 3:]]
            box = np.concatenate([np.concatenate(lo).min(0), np.concatenate(hi).max(0)])
        rec = bytearray(b"Obj_" + _str(o["name"]) + _u32(o.get("flags", 0)) + _u32(o.get("keyframe", NO_FRAME)) +
                        _u32(o.get("prev", NO_FRAME)) + np.asarray(box, "<f4").tobytes())
        uniq = sorted({int(m) for L in o["lods"] for key in ("mat_idx", "sphere_mat_idx")
                       for m in np.asarray(L.get(key, []), np.int64)})
        tail = (_u32(len(uniq)) + b"".join(struct.pack("<H", m) for m in uniq)) if version >= (1, 7) else b""
        obj_blobs.append((rec, lods, tail))
    header_len = 4 + 8 + 4 + 4 + 8 * len(objects)
    pos = objs_start + header_len
    obj_offsets, obj_bytes = [], []
    for rec, lods, tail in obj_blobs:
        obj_offsets.append(pos)
        lod_table_len = 4 + 8 * len(lods)
        lod_pos = pos + len(rec) + lod_table_len + len(tail)
        table = _u32(len(lods))
        for L in lods:
            table += _u64(lod_pos)
            lod_pos += len(L)
        blob = bytes(rec) + table + tail + b"".join(lods)
        obj_bytes.append(blob)
        pos += len(blob)
    inst_start = pos
    out += b"Objs" + _u64(inst_start) + _u32(flags) + _u32(len(objects)) + b"".join(_u64(o) for o in obj_offsets)
    for blob in obj_bytes:
        out += blob
    assert len(out) == inst_start
    out += b"Inst" + _u32(len(instances))
    for i in instances:
        m = np.asarray(i["world_to_inst"], "<f4").reshape(12)
        out += _str(i["name"]) + _u32(int(i["obj"])) + _u32(i.get("keyframe", NO_FRAME)) + _u32(NO_FRAME) + m.tobytes()
    with open(path, "wb") as f:
        f.write(bytes(out))


# ------------------------------------------------------------ scene export

_NDF_NAMES = {S.NDF_GGX: "GGX", S.NDF_BECKMANN: "BS"}
_SHADOWING_NAMES = {S.SHADOWING_VCAVITY: "vcavity", S.SHADOWING_SMITH: "smith"}


def _f(x):
    return float(np.float32(x))


def _material_json(m, texture_file):
    """JSON of a mufflon_amd material record; texture_file(id) -> file name."""
    out = {}
    if float(m["outerRefractionIndex"]) != 1.0 or np.any(np.asarray(m["outerAbsorption"]) != 0):
        out["outerMedium"] = {"ior": _f(m["outerRefractionIndex"]),
                              "absorption": [_f(x) for x in m["outerAbsorption"]]}

    def albedo():
        t = int(m["albedoTexture"])
        return texture_file(t) if t >= 0 else [_f(x) for x in m["albedo"]]

    def roughness():
        t = int(m["roughnessTexture"])
        if t >= 0:
            return texture_file(t)
        r = [_f(x) for x in m["roughness"]]
        return r[0] if r[0] == r[1] else r

    kind = int(m["type"])
    if kind == S.MAT_LAMBERT:
        out.update(type="lambert", albedo=albedo())
    elif kind == S.MAT_TORRANCE:
        out.update(type="torrance", albedo=albedo(), roughness=roughness(), ndf=_NDF_NAMES[int(m["ndf"])],
                   shadowingModel=_SHADOWING_NAMES[int(m["shadowing"])])
    elif kind in (S.MAT_WALTER, S.MAT_MICROFACET):
        out.update(type="walter" if kind == S.MAT_WALTER else "microfacet", roughness=roughness(), ndf=_NDF_NAMES[int(m["ndf"])],
                   shadowingModel=_SHADOWING_NAMES[int(m["shadowing"])], ior=_f(m["refractionIndex"]),
                   absorption=[_f(x) for x in m["absorption"]])
    elif kind == S.MAT_EMISSIVE:
        out.update(type="emissive", radiance=[_f(x) for x in m["radiance"]], scale=[_f(x) for x in m["scale"]])
    else:
        raise ValueError("material type %d has no scene-file form here" % kind)
    return out


def export_scene(scene, directory, name="scene", fov_degrees=None, background=None, deflate=False,
                 packed_normals=False, near=1e-4, far=1e10):
    """Write a one-LoD scene dict of mufflon_amd.scenes (verts, tris,
    optional quads / spheres, normals, uvs, mat_idx per primitive, materials, camera, optional textures,
    point_lights, envmap) as <name>.json + <name>.mff (+ PFM textures) in
    `directory`. fov_degrees defaults to the camera's vertical field of view;
    background: None, ("colored", rgb) or ("envmap", scale) with the scene's
    envmap texture. Returns the JSON path."""
    os.makedirs(directory, exist_ok=True)
    tex_files = {}

    def texture_file(i):
        if i not in tex_files:
            fn = "%s_tex%d.pfm" % (name, i)
            write_pfm(os.path.join(directory, fn), scene["textures"][i][0])
            tex_files[i] = fn
        return tex_files[i]

    mats = scene["materials"]
    bin_names = ["mat%d" % i for i in range(len(mats))]
    tris = np.asarray(scene["tris"]).reshape(-1, 3)
    quads = np.asarray(scene.get("quads", np.zeros((0, 4))), np.uint32).reshape(-1, 4)
    spheres = np.asarray(scene.get("spheres", np.zeros((0, 4))), np.float32).reshape(-1, 4)
    faces = len(tris) + len(quads)
    mat_idx = np.asarray(scene["mat_idx"])
    lod = dict(verts=scene["verts"], tris=tris, quads=quads, spheres=spheres, normals=scene["normals"],
               uvs=scene["uvs"], mat_idx=mat_idx[:faces], sphere_mat_idx=mat_idx[faces:])
    write_mff(os.path.join(directory, name + ".mff"), bin_names, [dict(name=name, lods=[lod])],
              [dict(name=name + "_inst", obj=0, world_to_inst=S.IDENTITY34)], deflate=deflate,
              packed_normals=packed_normals)
    cam = scene["camera"]
    if fov_degrees is None:
        fov_degrees = float(np.degrees(2.0 * np.arctan(float(cam["tan_vfov"]))))
    lights = {}
    points = scene.get("point_lights")
    for i, p in enumerate(points if points is not None else []):
        lights["point%d" % i] = {"type": "point", "position": [[_f(x) for x in p["position"]]],
                                 "intensity": [[_f(x) for x in p["intensity"]]]}
    if background is not None:
        if background[0] == "colored":
            lights["background"] = {"type": "envmap", "scale": [_f(x) for x in background[1]]}
        else:
            lights["background"] = {"type": "envmap", "map": texture_file(scene["envmap"]),
                                    "scale": [_f(x) for x in background[1]]}
    doc = {
        "version": "1.7",
        "binary": name + ".mff",
        "defaultScenario": "main",
        "cameras": {"camera": {"type": "pinhole", "path": [[_f(x) for x in cam["position"]]],
                               "viewDir": [[_f(x) for x in cam["view_dir"]]], "up": [[_f(x) for x in cam["up"]]],
                               "near": near, "far": far, "fov": fov_degrees}},
        "lights": lights,
        "materials": {b: _material_json(m, texture_file) for b, m in zip(bin_names, mats)},
        "scenarios": {"main": {"camera": "camera", "resolution": [int(cam["width"]), int(cam["height"])],
                               "lights": list(lights), "materialAssignments": {b: b for b in bin_names}}},
    }
    path = os.path.join(directory, name + ".json")
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)
    return path
