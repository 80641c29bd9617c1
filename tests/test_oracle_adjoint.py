"""The oracle's adjoint in its fixed-decision form (the derivative of one
realization of the estimator, every random decision and sampled direction
held) against central finite differences of the oracle's own renders with
identical seeds, texel by texel: the derivative of the L2 image loss with
respect to albedo texels (LAMBERT back wall, TORRANCE sphere), the emission
scale, and roughness texels (TORRANCE sphere, V-cavity and Smith). Roughness
differences use the replay render (the same paths, the same directions,
value * cos / pdf evaluated at the perturbed roughness). Differences in which
a decision flips are skipped. The default form, the derivative of the
expected image, is pinned by test_oracle_adjoint_expected."""
import numpy as np
import pytest

import pt_util as P
from mufflon_amd import scenes as S

SEED, DEPTH, NEE, ITS = 1, 6, 1, 2


def _loss(oracle, c, textures, materials=None):
    if materials is not None:
        c = dict(c, materials=materials)
    ps = P.oracle_pt_scene_textured(oracle, c, textures)
    img, _, _ = oracle.pt_render(ps, SEED, 0, DEPTH, NEE, ITS)
    return img.astype(np.float64)


def test_oracle_adjoint_matches_finite_differences(oracle):
    c = S.cornell_textured(8, 16, 24, 8, 8, sphere="albedo")
    target = _loss(oracle, c, [(S.procedural_texture(8, 8, 3, 0.15, 0.85, seed=11), s) for _, s in c["textures"]])
    img = _loss(oracle, c, c["textures"])
    delta = (2.0 * (img - target)).astype(np.float32)
    ps = P.oracle_pt_scene_textured(oracle, c)
    grads, em = oracle.pt_adjoint(ps, SEED, 0, DEPTH, NEE, ITS, delta, {0: (8, 8), 1: (8, 8)}, fixed_decisions=True)

    def loss_of(textures, materials=None):
        return float(np.sum((_loss(oracle, c, textures, materials) - target) ** 2))

    rng = np.random.default_rng(0)
    checked = agreed = 0
    for t in (0, 1):
        g = grads[t][..., :3]
        cand = np.argwhere(np.abs(g) > 1e-3 * np.abs(g).max())
        for k in rng.choice(len(cand), size=min(12, len(cand)), replace=False):
            y, x, ch = cand[k]
            fd = []
            for eps in (1e-3, 5e-4):
                tp = [(tx.copy(), s) for tx, s in c["textures"]]
                tm = [(tx.copy(), s) for tx, s in c["textures"]]
                tp[t][0][y, x, ch] += eps
                tm[t][0][y, x, ch] -= eps
                fd.append((loss_of(tp) - loss_of(tm)) / (2 * eps))
            if abs(fd[0] - fd[1]) > 1e-3 * abs(fd[0]):
                continue  # a random decision flipped inside the difference
            checked += 1
            agreed += abs(fd[0] - g[y, x, ch]) <= 1e-3 * abs(fd[0]) + 1e-9
    assert checked >= 16 and agreed >= checked - 1, (checked, agreed)
    # emission scale of the light (material 0), channel by channel
    for ch in range(3):
        eps = 1e-3
        mp = c["materials"].copy()
        mm = c["materials"].copy()
        mp["scale"][0, ch] += eps
        mm["scale"][0, ch] -= eps
        fd = (loss_of(c["textures"], mp) - loss_of(c["textures"], mm)) / (2 * eps)
        assert em[0, ch] == pytest.approx(fd, rel=1e-3)


@pytest.mark.parametrize("shadowing", [S.SHADOWING_VCAVITY, S.SHADOWING_SMITH])
def test_oracle_roughness_adjoint_matches_replay_fd(oracle, shadowing):
    c = S.cornell_textured(8, 16, 24, 8, 8, sphere="roughness", sphere_shadowing=shadowing)
    base = P.oracle_pt_scene_textured(oracle, c)
    img, _, _ = oracle.pt_render(base, SEED, 0, DEPTH, NEE, ITS)
    img = img.astype(np.float64)
    # the replay along the recorded directions reproduces the primal
    again = oracle.pt_render_replay(base, base, SEED, 0, DEPTH, NEE, ITS).astype(np.float64)
    np.testing.assert_allclose(again, img, rtol=1e-5, atol=1e-7)
    target = img * 0.8 + 0.05
    delta = (2.0 * (img - target)).astype(np.float32)
    grads, _ = oracle.pt_adjoint(base, SEED, 0, DEPTH, NEE, ITS, delta, {1: (8, 8)}, fixed_decisions=True)
    g = grads[1][..., 0]
    assert not grads[1][..., 1:].any()  # one channel drives both roughnesses

    def loss(rough):
        tex = [c["textures"][0], (rough, c["textures"][1][1])]
        ev = P.oracle_pt_scene_textured(oracle, c, tex)
        r = oracle.pt_render_replay(base, ev, SEED, 0, DEPTH, NEE, ITS).astype(np.float64)
        return float(np.sum((r - target) ** 2))

    rough0 = c["textures"][1][0]
    cand = np.argwhere(np.abs(g) > 1e-3 * np.abs(g).max())
    rng = np.random.default_rng(1)
    checked = agreed = 0
    for k in rng.choice(len(cand), size=min(16, len(cand)), replace=False):
        y, x = cand[k]
        fd = []
        for eps in (1e-3, 5e-4):
            rp, rm = rough0.copy(), rough0.copy()
            rp[y, x, 0] += eps
            rm[y, x, 0] -= eps
            fd.append((loss(rp) - loss(rm)) / (2 * eps))
        if abs(fd[0] - fd[1]) > 1e-3 * abs(fd[0]):
            continue  # a random decision flipped inside the difference
        checked += 1
        agreed += abs(fd[0] - g[y, x]) <= 1e-3 * abs(fd[0]) + 1e-9
    assert checked >= 10 and agreed >= checked - 1, (checked, agreed)
