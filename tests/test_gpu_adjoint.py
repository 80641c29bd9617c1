"""Adjoint (path replay) on the GPU.
* Against the oracle's FP64 adjoint on the same paths, in both forms (the
  derivative of the expected image, the default, and the fixed-decision
  derivative of one realization), for albedo and roughness textures with
  V-cavity and Smith shadowing.
* The differentiable benchmark's configuration (the Cornell box with a
  512x512 albedo texture on the back wall and a 256x256 roughness texture on
  the GGX sphere, 256x256 pixels, 4 paths per pixel):
  - fixed-decision form against central finite differences of GPU renders
    with identical seeds, texel by texel within 1e-3 relative (albedo
    texels; differences in which a random decision flips are counted and
    skipped; the roughness texels of this form are pinned per texel through
    the oracle, whose replay differences hold the sampled directions);
  - default form against central differences of the expected image over
    texel blocks (albedo and roughness), replicated over seeds, within 5
    standard errors.
* Through the torch.autograd binding."""
import numpy as np
import pytest

import pt_util as P
from mufflon_amd import hip
from mufflon_amd import scenes as S

pytestmark = pytest.mark.gpu


def _setup(hip_ctx, c, depth):
    pt, sc, ids = P.gpu_pt_textured(hip_ctx, c, max_path_length=depth)
    return pt, ids


@pytest.mark.parametrize("mode", [hip.ADJOINT_EXPECTED, hip.ADJOINT_FIXED_DECISIONS])
@pytest.mark.parametrize("sphere,shadowing", [("albedo", S.SHADOWING_VCAVITY), ("roughness", S.SHADOWING_VCAVITY),
                                              ("roughness", S.SHADOWING_SMITH)])
def test_adjoint_matches_oracle_fp64(hip_ctx, oracle, sphere, shadowing, mode):
    c = S.cornell_textured(12, 24, 40, 16, 16, sphere=sphere, sphere_shadowing=shadowing)
    print("adjoint vs oracle fp64, 40x40, sphere %s, shadowing %d, mode %d" % (sphere, shadowing, mode), flush=True)
    depth, its = 8, 3
    pt, ids = _setup(hip_ctx, c, depth)
    hip_ctx.pt_set_adjoint_mode(pt, mode)
    hip_ctx.pt_iterate(pt, its)
    img = hip_ctx.pt_target(pt)
    rng = np.random.default_rng(5)
    delta = rng.normal(size=img.shape).astype(np.float32)
    hip_ctx.pt_adjoint(pt, its, delta)
    shapes = [t.shape for t, _ in c["textures"]]
    g_gpu = [hip_ctx.pt_texture_gradient(pt, int(i), sh) for i, sh in zip(ids, shapes)]
    e_gpu = hip_ctx.pt_emission_gradient(pt, 0)
    ps = P.oracle_pt_scene_textured(oracle, c, light_tree=hip_ctx.pt_light_tree(pt))
    o_img, _, _ = oracle.pt_render(ps, 1, 0, depth, 1, its)
    assert np.array_equal(img.view(np.uint32), o_img.view(np.uint32))
    g_orc, e_orc = oracle.pt_adjoint(ps, 1, 0, depth, 1, its, delta, {0: (16, 16), 1: (16, 16)},
                                     fixed_decisions=mode == hip.ADJOINT_FIXED_DECISIONS)
    for t in (0, 1):
        ref = g_orc[t][..., :shapes[t][2]]
        scale = np.abs(ref).max()
        assert scale > 0
        err = np.abs(g_gpu[t] - ref)
        print("texture %d: max |grad| %.4g, max abs err %.3g" % (t, scale, float(err.max())), flush=True)
        assert (err <= 1e-3 * np.abs(ref) + 1e-5 * scale).all()
    np.testing.assert_allclose(e_gpu, e_orc[0], rtol=1e-3)


def test_adjoint_finite_differences_config3(hip_ctx):
    """Fixed-decision form: 256x256, 4 spp, L2 loss to a target rendered from
    a texture of seed 11; 64 random albedo texels with a non-negligible
    gradient."""
    c = S.cornell_textured(128, 256, 256, 512, 256)
    print("adjoint FD check, 256x256, 4 spp", flush=True)
    its, depth = 4, 8
    pt, ids = _setup(hip_ctx, c, depth)
    hip_ctx.pt_set_adjoint_mode(pt, hip.ADJOINT_FIXED_DECISIONS)
    textures = [t.copy() for t, _ in c["textures"]]
    target_tex = [S.procedural_texture(512, 512, 3, 0.15, 0.85, seed=11),
                  S.procedural_texture(256, 256, 1, 0.05, 0.8, seed=12)]

    def render(texs):
        for i, t in zip(ids, texs):
            hip_ctx.texture_update(int(i), t)
        hip_ctx.pt_reset(pt)
        hip_ctx.pt_iterate(pt, its)
        return hip_ctx.pt_target(pt).astype(np.float64)

    target = render(target_tex)
    img = render(textures)
    delta = (2.0 * (img - target)).astype(np.float32)
    hip_ctx.timer_begin()
    hip_ctx.pt_adjoint(pt, its, delta)
    ms = hip_ctx.timer_end()
    print("adjoint %d spp: %.2f ms" % (its, ms), flush=True)
    grads = [hip_ctx.pt_texture_gradient(pt, int(i), t.shape) for i, t in zip(ids, textures)]
    rng = np.random.default_rng(7)
    checked = agreed = flips = 0
    rels = []
    assert np.abs(grads[1]).max() > 0  # the roughness texture has a gradient
    for t in (0,):
        g = grads[t]
        cand = np.argwhere(np.abs(g) > 1e-2 * np.abs(g).max())
        for k in rng.choice(len(cand), size=min(64, len(cand)), replace=False):
            y, x, ch = cand[k]
            fds = []
            for eps in (1e-3, 5e-4):
                tp = [a.copy() for a in textures]
                tm = [a.copy() for a in textures]
                tp[t][y, x, ch] += eps
                tm[t][y, x, ch] -= eps
                fds.append((np.sum((render(tp) - target) ** 2) - np.sum((render(tm) - target) ** 2)) / (2 * eps))
            if abs(fds[0] - fds[1]) > 2e-3 * abs(fds[0]):
                flips += 1  # a roulette / light choice flipped inside the step
                continue
            checked += 1
            rel = abs(fds[0] - g[y, x, ch]) / abs(fds[0])
            rels.append(rel)
            agreed += rel <= 1e-3
    print("FD check: %d texels compared, %d within 1e-3 relative, %d skipped (discontinuity), median rel %.2e" %
          (checked, agreed, flips, float(np.median(rels))), flush=True)
    assert checked >= 48 and agreed >= 0.9 * checked


def test_adjoint_expected_block_fd_config3(hip_ctx):
    """Default form on the benchmark scene: for a fixed adjoint image delta
    (the L2 loss's at a first render), the gradient summed over a block of
    texels against the central difference of sum(delta * I) when the whole
    block moves by eps, both per seed over 8 seeds of 16 iterations; the mean
    difference must lie within 5 standard errors (plus an eps^2 allowance)."""
    c = S.cornell_textured(128, 256, 256, 512, 256)
    print("adjoint expected-image block FD, 256x256", flush=True)
    its, depth, seeds = 16, 8, range(1, 9)
    pt, ids = _setup(hip_ctx, c, depth)
    hip_ctx.pt_set_adjoint_mode(pt, hip.ADJOINT_EXPECTED)
    textures = [t.copy() for t, _ in c["textures"]]

    def render(texs, seed):
        for i, t in zip(ids, texs):
            hip_ctx.texture_update(int(i), t)
        hip_ctx.pt_set_params(pt, dict(seed=seed, max_path_length=depth, nee_count=1))
        hip_ctx.pt_iterate(pt, its)
        return hip_ctx.pt_target(pt).astype(np.float64)

    target = render([S.procedural_texture(512, 512, 3, 0.15, 0.85, seed=11),
                     S.procedural_texture(256, 256, 1, 0.05, 0.8, seed=12)], 99)
    delta = (2.0 * (render(textures, 99) - target)).astype(np.float32)
    # (texture, y0, x0, size, channel, eps): albedo blocks of the back wall, roughness blocks of the
    # sphere's camera-facing half (u < 0.5)
    blocks = [(0, 128, 128, 128, 0, 0.01), (0, 256, 256, 128, 1, 0.01), (0, 0, 384, 128, 2, 0.01),
              (1, 96, 32, 64, 0, 0.01), (1, 32, 64, 64, 0, 0.01), (1, 160, 48, 64, 0, 0.01)]
    pairs = {k: [] for k in range(len(blocks))}
    for seed in seeds:
        render(textures, seed)
        hip_ctx.pt_adjoint(pt, its, delta)
        grads = [hip_ctx.pt_texture_gradient(pt, int(i), t.shape) for i, t in zip(ids, textures)]
        for k, (t, y0, x0, n, ch, eps) in enumerate(blocks):
            g = float(grads[t][y0:y0 + n, x0:x0 + n, ch].sum())
            tp = [a.copy() for a in textures]
            tm = [a.copy() for a in textures]
            tp[t][y0:y0 + n, x0:x0 + n, ch] += eps
            tm[t][y0:y0 + n, x0:x0 + n, ch] -= eps
            fd = float(np.sum(delta * (render(tp, seed) - render(tm, seed)))) / (2 * eps)
            pairs[k].append((g, fd))
    for k, (t, y0, x0, n, ch, eps) in enumerate(blocks):
        a = np.array(pairs[k])
        d = a[:, 0] - a[:, 1]
        se = d.std(ddof=1) / np.sqrt(len(d))
        print("texture %d block (%d, %d, %d) ch %d: adjoint %.5g fd %.5g, difference %.3g +- %.3g" %
              (t, y0, x0, n, ch, a[:, 0].mean(), a[:, 1].mean(), d.mean(), se), flush=True)
        assert abs(d.mean()) <= 5.0 * se + 1e-3 * abs(a[:, 1].mean()), (k, d.mean(), se)


def test_torch_autograd_function(hip_ctx):
    torch = pytest.importorskip("torch")
    from mufflon_amd import torch_pt
    c = S.cornell_textured(8, 16, 32, 16, 16)
    pt, ids = _setup(hip_ctx, c, 6)
    tracer = torch_pt.Tracer(hip_ctx, pt, [int(i) for i in ids], [t.shape for t, _ in c["textures"]], 32, 32)
    texs = [torch.tensor(t, requires_grad=True) for t, _ in c["textures"]]
    target = torch.zeros(32, 32, 3)
    img = torch_pt.render(tracer, texs, 2)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    # the same through the C-ABI directly (float atomics: equal up to summation order)
    hip_ctx.pt_adjoint(pt, 2, (2.0 * (img.detach().numpy() - 0.0)).astype(np.float32))
    for i, t in zip(ids, texs):
        direct = hip_ctx.pt_texture_gradient(pt, int(i), tuple(t.shape))
        np.testing.assert_allclose(t.grad.numpy(), direct, rtol=1e-4, atol=1e-6 * np.abs(direct).max())
    assert float(texs[0].grad.abs().sum()) > 0
