"""The adjoint through the core C API (render_adjoint,
render_get_texture_gradient, render_get_emission_gradient,
render_set_texture_data, render_set_emission_scale): a textured Cornell box
built through the loader calls, its image and its albedo, roughness and
emission gradients against the oracle (image bit for bit within 1e-4,
gradients against the FP64 adjoint within 1e-3 relative), a texture change
against the oracle's render of the new texture, and the torch.autograd
binding over the core API against the direct calls."""
import ctypes as C
import os

import numpy as np
import pytest

import pt_util as P
from mufflon_amd import core as K
from mufflon_amd import scenes as S

pytestmark = pytest.mark.gpu

W = H = 40
DEPTH, ITS = 8, 3


def _tanf(x):
    libm = C.CDLL("libm.so.6")
    libm.tanf.restype = C.c_float
    libm.tanf.argtypes = [C.c_float]
    return np.float32(libm.tanf(np.float32(x)))


def _world(tmp_path, c):
    mi = K.CoreInterface()
    scen = K.build_world(mi, c, str(tmp_path), name="adj")
    assert mi.render_enable_renderer(0, 0)
    for name, v in (("Seed", 1), ("Max. path length", DEPTH), ("NEE count", 1)):
        assert mi.renderer_set_parameter_int(name, v)
    assert mi.world_load_scenario(scen), mi.core_get_dll_error()
    mi.render_enable_render_target("Radiance", False)
    texs = [mi.world_get_texture(os.path.join(str(tmp_path), "adj_tex%d.pfm" % i)) for i in range(len(c["textures"]))]
    assert all(texs)
    return mi, texs


def _oracle_scene(oracle, mi, c, textures=None):
    """Oracle scene with the camera as the core stores it (set_view_dir, and
    tanVFov = tanf(vFov / 2) of the float vFov it was given)."""
    cam = mi.world_get_camera("camera")
    d, u = K.Vec3(), K.Vec3()
    mi.lib.world_get_camera_direction(cam, C.byref(d), 0)
    mi.lib.world_get_camera_up(cam, C.byref(u), 0)
    vfov = np.float32(2.0 * np.arctan(float(c["camera"]["tan_vfov"])))
    c = dict(c)
    c["camera"] = dict(c["camera"], view_dir=np.array([d.x, d.y, d.z], np.float32),
                       up=np.array([u.x, u.y, u.z], np.float32), tan_vfov=_tanf(vfov / np.float32(2)))
    return P.oracle_pt_scene_textured(oracle, c, textures=textures)


def _stored(tex):
    """Texels in the core's stored format: RGB files become RGBA (alpha 0)."""
    if tex.shape[2] == 3:
        return np.concatenate([tex, np.zeros(tex.shape[:2] + (1,), np.float32)], axis=2)
    return tex


def _render(mi, its=ITS):
    assert mi.render_reset()
    for _ in range(its):
        ok, _ = mi.render_iterate()
        assert ok, mi.core_get_dll_error()
    return mi.mufflon_get_target_image("Radiance", False, W, H)


def _close(g, o, rtol=1e-4, atol=1e-6):
    return not (np.abs(g - o) > rtol * np.abs(o) + atol).any()


@pytest.mark.parametrize("sphere", ["roughness", "albedo"])
def test_core_adjoint_matches_oracle_fp64(oracle, tmp_path, sphere):
    print("core adjoint vs oracle fp64, sphere %s" % sphere, flush=True)
    c = S.cornell_textured(12, 24, W, 16, 16, sphere=sphere)
    mi, texs = _world(tmp_path, c)
    try:
        img = _render(mi)
        ps = _oracle_scene(oracle, mi, c)
        o_img, _, _ = oracle.pt_render(ps, 1, 0, DEPTH, 1, ITS)
        assert _close(img, o_img)
        delta = np.random.default_rng(3).normal(size=img.shape).astype(np.float32)
        mi.render_adjoint(delta)
        g_orc, e_orc = oracle.pt_adjoint(ps, 1, 0, DEPTH, 1, ITS, delta, {0: (16, 16), 1: (16, 16)})
        for k, (tex, _) in enumerate(c["textures"]):
            ch = _stored(tex).shape[2]
            g = mi.render_get_texture_gradient(texs[k], (16, 16, ch))
            used = tex.shape[2]
            ref = g_orc[k][..., :used]
            scale = np.abs(ref).max()
            err = np.abs(g[..., :used] - ref)
            print("texture %d: max |grad| %.4g, max abs err %.3g" % (k, scale, float(err.max())), flush=True)
            assert scale > 0
            assert (err <= 1e-3 * np.abs(ref) + 1e-5 * scale).all()
            assert not g[..., used:].any()  # the padding channel carries nothing
        e = mi.render_get_emission_gradient(0)
        np.testing.assert_allclose(e, e_orc[0], rtol=1e-3)
        # without a fresh iteration count the adjoint refuses (iteration -1)
        assert mi.render_reset()
        with pytest.raises(RuntimeError):
            mi.render_adjoint(delta)
    finally:
        mi.close()


def test_core_texture_data_and_emission_scale(oracle, tmp_path):
    c = S.cornell_textured(12, 24, W, 16, 16)
    mi, texs = _world(tmp_path, c)
    try:
        new = S.procedural_texture(16, 16, 3, 0.1, 0.9, seed=41)
        mi.render_set_texture_data(texs[0], _stored(new))
        img = _render(mi)
        ps = _oracle_scene(oracle, mi, c, textures=[(new, c["textures"][0][1]), c["textures"][1]])
        o_img, _, _ = oracle.pt_render(ps, 1, 0, DEPTH, 1, ITS)
        print("after render_set_texture_data: bit-exact %s" % np.array_equal(img.view(np.uint32),
                                                                           o_img.view(np.uint32)), flush=True)
        assert _close(img, o_img)
        # emission scale: the image of a light scaled by s is the image whose
        # light contributions are scaled by s; doubling the only light doubles
        # the whole image (nothing else emits)
        base = _render(mi)
        mat = c["materials"][0]
        mi.render_set_emission_scale(0, 2.0 * np.asarray(mat["scale"], np.float32))
        twice = _render(mi)
        np.testing.assert_allclose(twice, 2.0 * base, rtol=1e-5, atol=1e-7)
        with pytest.raises(RuntimeError):
            mi.render_set_emission_scale(1, (1.0, 1.0, 1.0))  # not EMISSIVE
    finally:
        mi.close()


def test_torch_over_core_api(tmp_path):
    torch = pytest.importorskip("torch")
    from mufflon_amd import torch_pt
    c = S.cornell_textured(8, 16, W, 16, 16)
    mi, texs = _world(tmp_path, c)
    try:
        shapes = [_stored(t).shape for t, _ in c["textures"]]
        tracer = torch_pt.CoreTracer(mi, W, H, texs, shapes, emission_slots=[0])
        params = [torch.tensor(_stored(t), requires_grad=True) for t, _ in c["textures"]]
        scale = torch.tensor(np.asarray(c["materials"][0]["scale"], np.float32), requires_grad=True)
        img = torch_pt.render(tracer, params, 2, emission_scales=[scale])
        loss = (img ** 2).sum()
        loss.backward()
        # the same through the core calls directly (float atomics: equal up to summation order)
        mi.render_adjoint((2.0 * img.detach().numpy()).astype(np.float32))
        for h, p, sh in zip(texs, params, shapes):
            direct = mi.render_get_texture_gradient(h, sh)
            np.testing.assert_allclose(p.grad.numpy(), direct, rtol=1e-4, atol=1e-6 * np.abs(direct).max())
        np.testing.assert_allclose(scale.grad.numpy(), mi.render_get_emission_gradient(0), rtol=1e-4)
        # the loss is quadratic in the emission scale (the image is linear in
        # it): dL/ds . s = 2 L
        assert float((scale.grad * scale.detach()).sum()) == pytest.approx(2.0 * float(loss.detach()), rel=1e-3)
        assert float(params[0].grad.abs().sum()) > 0
    finally:
        mi.close()
