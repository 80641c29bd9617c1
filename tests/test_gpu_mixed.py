"""Quads and spheres on the device against the oracle: the mixed-primitive
LBVH bit for bit (nodes and primitive ids), closest and any hit over mixed
scenes (one instance and an instance level; hit distances, ids and surface
parameters bit for bit), the path tracer over a room of bilinear quads,
triangle boxes and spheres with a quad light and a sphere light (every
target, megakernel and wavefront forms, instanced and not) and the wireframe
renderer over the same scene, and the adjoint (texture and emission
gradients) against the oracle's FP64 adjoint."""
import numpy as np
import pytest

import pt_util as P
import scene_util as U
from mufflon_amd import hip
from mufflon_amd import scenes as S
from mufflon_amd import workloads as W

pytestmark = pytest.mark.gpu

RTOL = 1e-4
ATOL = 1e-6


def _same(g, o, what):
    eq = np.array_equal(np.ascontiguousarray(g).view(np.uint32), np.ascontiguousarray(o).view(np.uint32))
    print("%s: bit-exact=%s, differing=%d" % (what, eq, int((g != o).sum())), flush=True)
    return eq


@pytest.mark.parametrize("nt,nq,ns,seed", [(3000, 2000, 1000, 1), (0, 4000, 0, 2), (0, 0, 5000, 3), (1, 1, 1, 4),
                                           (0, 1, 0, 5), (0, 0, 1, 6), (30000, 20000, 20000, 7)])
def test_mixed_lbvh_matches_oracle(hip_ctx, oracle, nt, nq, ns, seed):
    print("mixed LBVH: %d triangles, %d quads, %d spheres" % (nt, nq, ns), flush=True)
    v, t, q, sp = S.random_mixed_soup(nt, nq, ns, seed=seed)
    lod = hip_ctx.lod_create_mixed(v, t, q, sp)
    hip_ctx.lbvh_build(lod)
    nint, nodes, ids = hip_ctx.lbvh_get(lod, nt + nq + ns)
    b = oracle.lbvh_build_mixed(v, t, q, sp)
    assert nint == b["num_internal"]
    assert np.array_equal(ids, b["sorted_ids"])
    assert np.array_equal(nodes, b["nodes"])


@pytest.mark.parametrize("instances", [1, 16])
def test_mixed_trace_matches_oracle(hip_ctx, oracle, instances):
    print("mixed trace: %d instances" % instances, flush=True)
    rng = np.random.default_rng(20 + instances)
    geoms = [S.random_mixed_soup(nt, nq, ns, seed=30 + k, mu=-2.5) for k, (nt, nq, ns) in
             enumerate(((4000, 3000, 2000), (0, 600, 0), (0, 0, 800), (500, 0, 1), (0, 0, 1)))]
    if instances == 1:
        inst_lod, w = np.array([0], np.int32), S.IDENTITY34.reshape(1, 12)
        rays = U.rays_toward(rng, 20000, (-1, 2), (0, 1))
    else:
        inst_lod = rng.integers(0, len(geoms), instances).astype(np.int32)
        w = U.random_transforms(instances, seed=40, spread=2.0)
        rays = U.rays_toward(rng, 20000)
    lods = []
    for v, t, q, sp in geoms:
        lod = hip_ctx.lod_create_mixed(v, t, q, sp)
        hip_ctx.lbvh_build(lod)
        lods.append(lod)
    sc = hip_ctx.scene_create(np.array(lods, np.int32)[inst_lod], w)
    gt, gi, gp, gb, st = hip_ctx.trace_closest(sc, rays)
    scene, _ = U.oracle_scene(oracle, geoms, inst_lod, w)
    ot, oi, op, ob = oracle.trace_closest(scene, rays)
    assert (st == 0).all()
    hit = oi >= 0
    print("hits: %d of %d" % (int(hit.sum()), len(rays)), flush=True)
    assert hit.sum() > 1000
    assert np.array_equal(gi, oi) and np.array_equal(gp, op)
    assert _same(gt, ot, "t")
    assert _same(gb[hit], ob[hit], "surface parameters")
    assert np.array_equal(hip_ctx.trace_any(sc, rays), oracle.trace_any(scene, rays))


@pytest.fixture(scope="module", params=[True, False], ids=["instanced", "single"])
def mixed(request, hip_ctx, oracle):
    sc = S.mixed_scene(48, 40, instanced=request.param)
    pt, scid, _ = W.path_tracer(hip_ctx, sc, max_path_length=8)
    for t in range(8):
        hip_ctx.pt_enable_target(pt, t, True)
    box, _, _, _ = hip_ctx.scene_get(scid, len(sc["inst_lod"]))
    ps = P.oracle_pt_scene_general(oracle, sc, hip_ctx.pt_light_tree(pt), box)
    ref, nans = oracle.pt_render_targets(ps, 1, 0, 8, 1, 3, list(range(8)))
    return request.param, pt, ps, ref, nans


@pytest.mark.parametrize("form,tail", [("megakernel", 0), ("wavefront", 0), ("wavefront", 1)])
def test_mixed_pt_matches_oracle(hip_ctx, oracle, mixed, form, tail):
    instanced, pt, ps, ref, nans = mixed
    print("mixed pt: instanced %s, %s, tail after %d" % (instanced, form, tail), flush=True)
    hip_ctx.pt_set_tuning(pt, tail_after=tail)
    hip_ctx.pt_set_mode(pt, hip.PT_MEGAKERNEL if form == "megakernel" else hip.PT_WAVEFRONT, form == "wavefront")
    hip_ctx.pt_reset(pt)
    hip_ctx.pt_iterate(pt, 3)
    g_hit = hip_ctx.pt_target(pt, hip.TARGET_HIT_ID)
    assert np.array_equal(g_hit, ref[hip.TARGET_HIT_ID])
    # every primitive kind is seen
    h = ref[hip.TARGET_HIT_ID][..., 1]
    assert (h >= 24).any() and (h >= 30).any()
    for t in range(8):
        g, o = hip_ctx.pt_target(pt, t), ref[t]
        bad = np.abs(g - o) > RTOL * np.abs(o) + ATOL
        exact = _same(g, o, "target %d" % t)
        assert not bad.any(), t
        assert exact, t
    assert hip_ctx.pt_nan_count(pt) == nans
    assert ref[hip.TARGET_RADIANCE].mean() > 0.01


def test_mixed_wireframe_matches_oracle(hip_ctx, oracle, mixed):
    instanced, pt, ps, _, _ = mixed
    print("mixed wireframe: instanced %s" % instanced, flush=True)
    hip_ctx.wireframe_reset(pt, variance=True)
    hip_ctx.wireframe_iterate(pt, 2, line_width=1.5, max_trace_depth=64)
    o, ov, _ = oracle.wireframe_render(ps, 2, line_width=1.5, max_trace_depth=64, variance=True)
    g = hip_ctx.wireframe_border(pt)
    assert _same(g, o, "border")
    assert _same(hip_ctx.wireframe_border(pt, variance=True), ov, "border variance")
    assert (g == 1.0).any()


def test_mixed_scene_is_refused_by_a_batch(hip_ctx):
    sc = S.mixed_scene(16, 16, instanced=False)
    pt, _, _ = W.path_tracer(hip_ctx, sc, max_path_length=4)
    with pytest.raises(hip.HipError):
        hip_ctx.batch_create([pt])


@pytest.mark.parametrize("mode", [hip.ADJOINT_EXPECTED, hip.ADJOINT_FIXED_DECISIONS])
def test_mixed_adjoint_matches_oracle_fp64(hip_ctx, oracle, mode):
    """The adjoint over quads and spheres: an albedo texture on the white
    quads and triangle box (bilinear-patch uvs), a roughness texture on the
    GGX box and spheres (spherical uvs), emission gradients of the quad and
    sphere lights, against the oracle's FP64 adjoint of the same paths."""
    print("mixed adjoint vs oracle fp64, mode %d" % mode, flush=True)
    sc = S.mixed_scene(40, 32, instanced=True)
    mats = sc["materials"].copy()
    WHITE, GGX = 1, 4
    mats["albedoTexture"][WHITE] = 0
    mats["roughnessTexture"][GGX] = 1
    sc = dict(sc, materials=mats,
              textures=[(S.procedural_texture(16, 16, 3, 0.3, 0.9, seed=5), S.SAMPLING_LINEAR),
                        (S.procedural_texture(16, 16, 1, 0.05, 0.6, seed=6), S.SAMPLING_LINEAR)])
    depth, its = 6, 3
    pt, scid, tex = W.path_tracer(hip_ctx, sc, max_path_length=depth)
    hip_ctx.pt_set_adjoint_mode(pt, mode)
    hip_ctx.pt_iterate(pt, its)
    img = hip_ctx.pt_target(pt)
    delta = np.random.default_rng(7).normal(size=img.shape).astype(np.float32)
    hip_ctx.pt_adjoint(pt, its, delta)
    g_gpu = [hip_ctx.pt_texture_gradient(pt, int(i), t.shape) for i, (t, _) in zip(tex, sc["textures"])]
    lights = [0, 7]  # the quad light's and the sphere light's materials
    e_gpu = np.stack([hip_ctx.pt_emission_gradient(pt, m) for m in lights])
    box, _, _, _ = hip_ctx.scene_get(scid, len(sc["inst_lod"]))
    ps = P.oracle_pt_scene_general(oracle, sc, hip_ctx.pt_light_tree(pt), box)
    o_img, _, _ = oracle.pt_render(ps, 1, 0, depth, 1, its)
    assert _same(img, o_img, "primal")
    g_orc, e_orc = oracle.pt_adjoint(ps, 1, 0, depth, 1, its, delta, {0: (16, 16), 1: (16, 16)},
                                     fixed_decisions=mode == hip.ADJOINT_FIXED_DECISIONS)
    for t, ch in ((0, 3), (1, 1)):
        ref = g_orc[t][..., :ch]
        scale = np.abs(ref).max()
        err = np.abs(g_gpu[t] - ref)
        print("texture %d: max |grad| %.4g, max abs err %.3g" % (t, scale, float(err.max())), flush=True)
        assert scale > 0
        assert (err <= 1e-3 * np.abs(ref) + 1e-5 * scale).all()
    assert (np.abs(e_orc[lights]) > 0).all()
    # float atomics against double sums: relative to each value, with a floor
    # for the channels in which large terms cancel
    np.testing.assert_allclose(e_gpu, e_orc[lights], rtol=1e-3, atol=1e-4 * np.abs(e_orc[lights]).max())
