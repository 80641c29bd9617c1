"""The oracle's adjoint in its default form (the derivative of the expected
image) against derivatives known without the path tracer: a plane under a
uniform background (white furnace), whose expected radiance is
Le * a(theta_o), a being the plane BSDF's directional albedo at the viewing
angle.
* Lambert with a 1x1 albedo texture rho: a = rho, so d E[L] / d rho = Le per
  pixel and channel, exactly.
* Torrance (GGX, V-cavity or Smith) with a 1x1 roughness texture r:
  a(theta_o; r) = integral of value * cos over the hemisphere, evaluated by
  Gauss-Legendre x trapezoid quadrature of the oracle's own BSDF (pinned
  separately in test_oracle_materials) at r +- h, and integrated over each
  pixel's footprint.
The adjoint runs on independent seeds; the mean of the replicates must lie
within 5 standard errors of the closed form. The fixed-decision form (the
derivative of one realization, roulette and sampled directions held) is
shown to be biased on the same scenes.
"""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from closed_form_util import GAUSS3, camera, oracle_scene, pixel_points, plane_scene
from mufflon_amd import hip
from mufflon_amd import scenes as S

LE = np.array([0.3, 0.4, 0.6]) * 0.8
SIZE = 16


def _furnace(O, plane_material, texture):
    cam = camera(SIZE)
    c, _ = plane_scene(plane_material=plane_material)
    bg = hip.background_record(hip.BACKGROUND_COLORED, color=(0.3, 0.4, 0.6), scale=0.8)
    return oracle_scene(O, c, cam, background=bg, textures=[(texture, S.SAMPLING_NEAREST)]), cam


def _replicates(O, ps, its, seeds, fixed):
    """Adjoint of sum(Radiance) (delta = 1) for several seeds: texel 0 of
    texture 0 per replicate, (n, 3)."""
    delta = np.ones((SIZE, SIZE, 3), np.float32)
    out = []
    for s in seeds:
        g, _ = O.pt_adjoint(ps, s, 0, 2, 1, its, delta, {0: (1, 1)}, fixed_decisions=fixed)
        out.append(g[0][0, 0, :3])
    return np.array(out)


def _agrees(reps, expected):
    m = reps.mean(0)
    se = reps.std(0, ddof=1) / np.sqrt(len(reps))
    return bool((np.abs(m - expected) <= 5.0 * se + 1e-9).all()), m, se


def _albedo_case(O):
    rho = (0.55, 0.35, 0.2)
    mat = S.material(S.MAT_LAMBERT, albedo_texture=0)
    tex = np.array(rho, np.float32).reshape(1, 1, 3)
    ps, _ = _furnace(O, mat, tex)
    return ps, LE * SIZE * SIZE


def test_albedo_gradient_furnace(oracle):
    ps, expected = _albedo_case(oracle)
    ok, m, se = _agrees(_replicates(oracle, ps, 16, range(1, 9), False), expected)
    assert ok, (m, expected, se)


def test_fixed_decision_albedo_gradient_is_biased(oracle):
    """The roulette's continuation probability depends on the albedo: the
    derivative with roulette decisions held misses the flips' share."""
    ps, expected = _albedo_case(oracle)
    reps = _replicates(oracle, ps, 16, range(1, 9), True)
    m = reps.mean(0)
    se = reps.std(0, ddof=1) / np.sqrt(len(reps))
    assert (np.abs(m - expected) > 10.0 * se).any(), (m, expected, se)


def directional_albedo_derivative(O, material_of, thetas, r, h=0.005, n_mu=64, n_phi=128):
    """d/dr of a(theta_o; r) = int value(o, i) cos_i di (3 channels) at the
    viewing angles thetas, by quadrature of the oracle's BSDF at r +- h."""
    mu, wmu = np.polynomial.legendre.leggauss(n_mu)
    mu = 0.5 * (mu + 1.0)
    wmu = 0.5 * wmu
    phi = (np.arange(n_phi) + 0.5) * (2.0 * np.pi / n_phi)
    M, P = np.meshgrid(mu, phi, indexing="ij")
    st = np.sqrt(1.0 - M * M)
    ex = np.stack([st * np.cos(P), st * np.sin(P), M], -1).reshape(-1, 3)
    w = (np.outer(wmu, np.full(n_phi, 2.0 * np.pi / n_phi)) * M).reshape(-1)
    out = np.zeros((len(thetas), 3))
    for sign in (1.0, -1.0):
        mat = material_of(r + sign * h)
        for k, th in enumerate(thetas):
            inn = np.broadcast_to([np.sin(th), 0.0, np.cos(th)], ex.shape)
            val, _, _ = O.bsdf_evaluate_batch(mat, inn, ex)
            out[k] += sign * (val.astype(np.float64) * w[:, None]).sum(0)
    return out / (2.0 * h)


def _roughness_case(O, shadowing):
    albedo, r = (0.8, 0.6, 0.4), 0.45

    def material_of(rr):
        return S.material(S.MAT_TORRANCE, albedo=albedo, roughness=rr, shadowing=shadowing)

    mat = S.material(S.MAT_TORRANCE, albedo=albedo, roughness_texture=0, shadowing=shadowing)
    ps, cam = _furnace(O, mat, np.full((1, 1, 1), r, np.float32))
    # viewing angle of every quadrature node of every pixel
    x, wq = pixel_points(cam, *GAUSS3)
    o = cam["position"].astype(np.float64) - x
    theta = np.arccos(o[..., 1] / np.linalg.norm(o, axis=-1))
    grid = np.linspace(theta.min() - 0.01, theta.max() + 0.01, 40)
    da = CubicSpline(grid, directional_albedo_derivative(O, material_of, grid, r))(theta)
    # one texture channel drives both roughnesses; delta = 1 sums the image's channels
    expected = float(((da * wq[None, None, :, None]).sum(2) * LE).sum())
    return ps, expected


@pytest.mark.parametrize("shadowing", [S.SHADOWING_VCAVITY, S.SHADOWING_SMITH])
def test_roughness_gradient_furnace(oracle, shadowing):
    ps, expected = _roughness_case(oracle, shadowing)
    reps = _replicates(oracle, ps, 64, range(1, 9), False)[:, :1]
    ok, m, se = _agrees(reps, np.array([expected]))
    assert ok, (m, expected, se)


@pytest.mark.parametrize("shadowing", [S.SHADOWING_VCAVITY, S.SHADOWING_SMITH])
def test_fixed_decision_roughness_gradient_is_biased(oracle, shadowing):
    """Holding the sampled directions fixed drops the sampling's share of the
    derivative (for V-cavity the walk term d ln value - d ln pdf is zero)."""
    ps, expected = _roughness_case(oracle, shadowing)
    reps = _replicates(oracle, ps, 16, range(1, 9), True)[:, 0]
    assert abs(reps.mean() - expected) > 10.0 * reps.std(ddof=1) / np.sqrt(len(reps)), (reps.mean(), expected)
