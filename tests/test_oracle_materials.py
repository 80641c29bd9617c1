"""Pins of the oracle's materials, textures and background sampler against
closed forms and against each other (no device needed):

* GGX / Beckmann NDFs integrate to 1 over projected solid angle;
* the dielectric Fresnel term matches the textbook unpolarised formula, its
  normal-incidence value ((n-1)/(n+1))^2 and total internal reflection;
* every sampler's reported pdf and throughput agree with the evaluator at the
  sampled direction, and the sampled distribution integrates the evaluator
  (Monte Carlo mean of the throughput = quadrature of f cos);
* texture sampling follows the CPU texture conventions (texel corners, no
  half-texel offset, wrap-around);
* the environment-map sampler's pdf equals the evaluator's pdf and integrates
  to 1; the colored background's flux has its closed form.
"""
import numpy as np
import pytest

from mufflon_amd import hip
from mufflon_amd import scenes as S


def _sphere_grid(n_theta, n_phi, upper=True):
    """Midpoint quadrature over the (upper) sphere: directions and weights dw."""
    th_max = np.pi / 2 if upper else np.pi
    th = (np.arange(n_theta) + 0.5) * th_max / n_theta
    ph = (np.arange(n_phi) + 0.5) * 2 * np.pi / n_phi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1)
    w = np.sin(T) * (th_max / n_theta) * (2 * np.pi / n_phi)
    return d.reshape(-1, 3), w.reshape(-1)


@pytest.mark.parametrize("ndf,rx,ry", [(S.NDF_GGX, 0.3, 0.3), (S.NDF_GGX, 0.15, 0.5), (S.NDF_BECKMANN, 0.3, 0.3),
                                       (S.NDF_BECKMANN, 0.2, 0.45)])
def test_ndf_normalisation(oracle, ndf, rx, ry):
    d, w = _sphere_grid(400, 200)
    D = np.array([oracle.eval_ndf(ndf, rx, ry, h) for h in d], np.float64)
    integral = float(np.sum(D * d[:, 2] * w))
    assert abs(integral - 1.0) < 5e-3


def test_fresnel_closed_forms(oracle):
    for n in (1.33, 1.5, 2.4):
        f, ct = oracle.fresnel_dielectric(1.0, n, 1.0)
        assert abs(f - ((n - 1) / (n + 1)) ** 2) < 1e-6 and ct == 1.0
    rng = np.random.default_rng(3)
    for _ in range(200):
        n_i, n_t = rng.choice([1.0, 1.5]), rng.choice([1.0, 1.33, 1.5])
        ci = float(rng.uniform(0.01, 1.0))
        f, ct = oracle.fresnel_dielectric(n_i, n_t, ci)
        s2 = (n_i / n_t) ** 2 * (1 - ci * ci)
        if s2 >= 1:
            assert f == pytest.approx(1.0, abs=1e-6) and ct == 0.0  # total internal reflection
            continue
        c_t = np.sqrt(1 - s2)
        rs = (n_i * ci - n_t * c_t) / (n_i * ci + n_t * c_t)
        rp = (n_t * ci - n_i * c_t) / (n_t * ci + n_i * c_t)
        assert f == pytest.approx(0.5 * (rs * rs + rp * rp), rel=2e-5, abs=1e-7)
        assert ct == pytest.approx(c_t, rel=1e-5)


MATS = {
    "lambert": S.material(S.MAT_LAMBERT, albedo=(0.8, 0.6, 0.4)),
    "torrance_ggx_vcavity": S.material(S.MAT_TORRANCE, albedo=(0.9, 0.7, 0.5), roughness=0.3),
    "torrance_ggx_aniso": S.material(S.MAT_TORRANCE, albedo=(1, 1, 1), roughness=(0.15, 0.45)),
    "torrance_beckmann_vcavity": S.material(S.MAT_TORRANCE, albedo=(1, 1, 1), roughness=0.25, ndf=S.NDF_BECKMANN),
    "torrance_ggx_smith": S.material(S.MAT_TORRANCE, albedo=(1, 1, 1), roughness=0.35,
                                     shadowing=S.SHADOWING_SMITH),
    "microfacet_ggx": S.material(S.MAT_MICROFACET, roughness=0.2, ior=1.5),
    "microfacet_ggx_smith": S.material(S.MAT_MICROFACET, roughness=0.3, ior=1.5, shadowing=S.SHADOWING_SMITH),
    "walter_ggx": S.material(S.MAT_WALTER, roughness=0.2, ior=1.5),
    "walter_ggx_smith": S.material(S.MAT_WALTER, roughness=0.3, ior=1.33, shadowing=S.SHADOWING_SMITH),
    "walter_beckmann": S.material(S.MAT_WALTER, roughness=0.25, ior=1.5, ndf=S.NDF_BECKMANN),
}


def _incident(theta, phi, below=False):
    # incident points towards the surface: tangent-space incidentTS has z = cos > 0 from above
    d = np.array([-np.sin(theta) * np.cos(phi), -np.sin(theta) * np.sin(phi), np.cos(theta)], np.float32)
    if below:
        d[2] = -d[2]
    return d


@pytest.mark.parametrize("name", sorted(MATS))
def test_sample_matches_evaluate(oracle, name):
    """At every sampled direction the evaluator reports the sampler's pdf and
    value * |cos| / pdf = the sampler's throughput."""
    m = MATS[name]
    rng = np.random.default_rng(11)
    checked = 0
    for k in range(400):
        inc = _incident(rng.uniform(0.05, 1.4), rng.uniform(0, 2 * np.pi), below=(k % 4 == 3))
        u0, u1 = rng.random(2).astype(np.float32)
        i0 = int(rng.integers(0, 2 ** 63)) * 2 + int(rng.integers(0, 2))
        valid, tp, ex, pf, pb = oracle.bsdf_sample(m, inc, float(u0), float(u1), i0)
        if not valid or pf <= 1e-3 or pf > 1e6:
            continue
        val, ef, eb = oracle.bsdf_evaluate(m, inc, ex)
        assert ef == pytest.approx(pf, rel=2e-3), (name, k)
        assert eb == pytest.approx(pb, rel=2e-3), (name, k)
        np.testing.assert_allclose(val * abs(ex[2]) / ef, tp, rtol=3e-3, atol=1e-6)
        checked += 1
    assert checked > 200


@pytest.mark.parametrize("name", ["lambert", "torrance_ggx_vcavity", "torrance_ggx_smith", "torrance_beckmann_vcavity",
                                  "microfacet_ggx", "walter_ggx", "walter_ggx_smith"])
def test_sampled_distribution_integrates_the_bsdf(oracle, name):
    """E[throughput] over the sampler = integral of f |cos| over the sphere."""
    m = MATS[name]
    inc = _incident(0.7, 0.3)
    rng = np.random.default_rng(5)
    n = 40000
    u = rng.random((n, 2)).astype(np.float32)
    ii = rng.integers(0, 2 ** 62, n) * 4 + rng.integers(0, 4, n)
    acc = np.zeros(3)
    for k in range(n):
        valid, tp, _, _, _ = oracle.bsdf_sample(m, inc, float(u[k, 0]), float(u[k, 1]), int(ii[k]))
        if valid:
            acc += tp
    mc = acc / n
    d, w = _sphere_grid(240, 240, upper=False)
    f = np.array([oracle.bsdf_evaluate(m, inc, e)[0][0] for e in d.astype(np.float32)])
    quad = float(np.sum(f * np.abs(d[:, 2]) * w))
    assert mc[0] == pytest.approx(quad, rel=0.03, abs=2e-3)


def _snell(inc_ts, eta):
    """Refracted direction of an incidentTS (the direction towards where the
    path came from, z > 0) about the normal (0, 0, 1), n_i / n_t = eta;
    None beyond the critical angle."""
    c = inc_ts[2]
    k = 1.0 - eta * eta * (1.0 - c * c)
    if k < 0:
        return None
    return -eta * inc_ts + (eta * c - np.sqrt(k)) * np.array([0, 0, 1.0])


@pytest.mark.parametrize("ndf", [S.NDF_GGX, S.NDF_BECKMANN])
def test_walter_smooth_limit_is_snell_refraction(oracle, ndf):
    """At vanishing roughness a WALTER sample is Snell's refraction of the
    incident direction about the normal (the half vector is the normal),
    with throughput 1: no Fresnel term, no reflection below the critical
    angle."""
    m = S.material(S.MAT_WALTER, roughness=1e-4, ior=1.5, ndf=ndf)
    rng = np.random.default_rng(8)
    for k in range(200):
        th = rng.uniform(0.0, 1.45)
        inc = _incident(th, rng.uniform(0, 2 * np.pi))
        u0, u1 = (float(x) for x in rng.random(2).astype(np.float32))
        valid, tp, ex, pf, _ = oracle.bsdf_sample(m, inc, u0, u1, int(rng.integers(0, 2 ** 63)))
        assert valid
        expect = _snell(inc.astype(np.float64), 1.0 / 1.5)
        np.testing.assert_allclose(ex, expect, atol=2e-3)
        np.testing.assert_allclose(tp, 1.0, atol=2e-3)


def test_walter_total_internal_reflection(oracle):
    """From inside (below, n_i = 1.5 to n_t = 1) beyond the critical angle
    asin(1 / 1.5) every smooth sample is the mirror direction; evaluating a
    reflection below the critical angle gives 0, and the transmitted
    direction of a sample from inside evaluates to the sample's pdf."""
    m = S.material(S.MAT_WALTER, roughness=1e-4, ior=1.5)
    crit = np.arcsin(1 / 1.5)
    rng = np.random.default_rng(12)
    for k in range(100):
        th = rng.uniform(crit + 0.05, 1.45)
        phi = rng.uniform(0, 2 * np.pi)
        inc = _incident(th, phi, below=True)
        valid, tp, ex, pf, _ = oracle.bsdf_sample(m, inc, *(float(x) for x in rng.random(2).astype(np.float32)),
                                                  int(rng.integers(0, 2 ** 63)))
        assert valid
        mirror = np.array([-inc[0], -inc[1], inc[2]], np.float64)  # the mirror stays on the incident's side
        np.testing.assert_allclose(ex, mirror, atol=2e-3)
        np.testing.assert_allclose(tp, 1.0, atol=2e-3)
    # no reflection outside total internal reflection (rough, from above)
    m = S.material(S.MAT_WALTER, roughness=0.3, ior=1.5)
    inc = _incident(0.4, 0.2)
    refl = np.array([-inc[0], -inc[1], inc[2]], np.float32)
    val, pf, pb = oracle.bsdf_evaluate(m, inc, refl)
    assert not np.any(val) and pf == 0.0


def test_walter_transmits_what_microfacet_does_not_reflect(oracle):
    """For the same roughness and ior, the MICROFACET material's transmitted
    lobe is WALTER's lobe weighted by 1 - Fresnel of the half vector: on any
    refracted pair their values differ by exactly that factor."""
    w = S.material(S.MAT_WALTER, roughness=0.3, ior=1.5)
    mf = S.material(S.MAT_MICROFACET, roughness=0.3, ior=1.5)
    rng = np.random.default_rng(14)
    checked = 0
    for k in range(300):
        inc = _incident(rng.uniform(0.05, 1.3), rng.uniform(0, 2 * np.pi))
        valid, tp, ex, pf, _ = oracle.bsdf_sample(w, inc, *(float(x) for x in rng.random(2).astype(np.float32)),
                                                  int(rng.integers(0, 2 ** 63)))
        if not valid or ex[2] > 0:
            continue
        vw, pw, _ = oracle.bsdf_evaluate(w, inc, ex)
        vm, pm, _ = oracle.bsdf_evaluate(mf, inc, ex)
        if vw[0] <= 0:
            continue
        # half vector of the refracted pair (n_i = 1, n_t = 1.5), facing the incident side
        i = inc.astype(np.float64)
        h = 1.0 * i + 1.5 * ex.astype(np.float64)
        h /= np.linalg.norm(h) * np.sign(h[2])
        f, _ = oracle.fresnel_dielectric(1.0, 1.5, abs(float(i @ h)))
        assert vm[0] == pytest.approx((1 - f) * vw[0], rel=2e-4)
        assert pm == pytest.approx((1 - f) * pw, rel=2e-4)
        checked += 1
    assert checked > 100


def test_torrance_reciprocity(oracle):
    m = MATS["torrance_ggx_aniso"]
    rng = np.random.default_rng(2)
    for _ in range(100):
        a, b = _incident(rng.uniform(0.1, 1.4), rng.uniform(0, 6.28)), _incident(rng.uniform(0.1, 1.4),
                                                                                  rng.uniform(0, 6.28))
        ex = -b  # excident of the swapped pair
        v1 = oracle.bsdf_evaluate(m, a, ex)[0]
        v2 = oracle.bsdf_evaluate(m, b, -a)[0]
        np.testing.assert_allclose(v1, v2, rtol=2e-4, atol=1e-7)


def test_texture_conventions(oracle):
    W, H = 8, 4
    ramp = np.zeros((H, W, 3), np.float32)
    ramp[:, :, 0] = np.arange(W)[None, :]
    ramp[:, :, 1] = np.arange(H)[:, None]
    ramp[:, :, 2] = 7.0
    tx = oracle.Textures([(ramp, 0), (ramp, 1), (np.full((1, 1, 3), 0.3, np.float32), 1)])
    # nearest: texel floor(uv * size), wrap-around for negative / >1 coordinates
    assert np.array_equal(oracle.texture_sample(tx, 0, 0.49, 0.5)[:3], [3, 2, 7])
    assert np.array_equal(oracle.texture_sample(tx, 0, -0.01, 1.01)[:3], [7, 0, 7])
    # linear: exact at texel corners (no half-texel offset) ...
    for x in range(W):
        assert np.array_equal(oracle.texture_sample(tx, 1, x / W, 1 / H)[:3], [x, 1, 7])
    # ... linear in between, and the last column blends towards the wrapped first one
    s = oracle.texture_sample(tx, 1, 2.25 / W, 1.5 / H)
    assert s[0] == pytest.approx(2.25, abs=1e-6) and s[1] == pytest.approx(1.5, abs=1e-6)
    s = oracle.texture_sample(tx, 1, 7.5 / W, 0.0)
    assert s[0] == pytest.approx(3.5, abs=1e-6)  # 0.5 * 7 + 0.5 * 0
    # a 1x1 texture returns its value exactly
    assert np.array_equal(oracle.texture_sample(tx, 2, 0.37, 0.91)[:3], np.float32([0.3, 0.3, 0.3]))


def _env(h=32, w=64, seed=4):
    rng = np.random.default_rng(seed)
    env = rng.uniform(0.1, 1.0, (h, w, 3)).astype(np.float32)
    env[h // 3:h // 3 + 3, w // 2:w // 2 + 4] = 40.0  # a bright patch
    return env


def test_envmap_pdf_consistency(oracle):
    env = _env()
    box = np.array([-1, 0, -1, 1, 2, 1], np.float32)
    bg = hip.background_record(hip.BACKGROUND_ENVMAP, scale=1.5, envmap_texture=0)
    flux, sat = hip.background_build(bg, env, box)
    tx = oracle.Textures([(env, hip.SAMPLING_NEAREST)])
    bs = oracle.background_scene(bg[0], tx, box, sat=sat, flux=flux)
    rng = np.random.default_rng(9)
    checked = 0
    for _ in range(500):
        u0, u1 = (float(x) for x in rng.random(2).astype(np.float32))
        d, pdf, rad = oracle.background_sample(bs, u0, u1)
        assert abs(np.linalg.norm(d) - 1) < 1e-5
        rad_e, pdf_e = oracle.background_evaluate(bs, d)
        # away from texel borders the evaluation lands in the sampled texel
        uu = (np.arctan2(d[2], d[0]) / (2 * np.pi)) % 1.0 * env.shape[1]
        vv = (1 - np.arccos(np.clip(d[1], -1, 1)) / np.pi) * env.shape[0]
        if min(abs(uu - round(uu)), abs(vv - round(vv))) < 0.02:
            continue
        np.testing.assert_allclose(rad_e, rad, rtol=1e-6)
        assert pdf_e == pytest.approx(pdf, rel=2e-3)
        checked += 1
    assert checked > 400
    # the evaluator's pdf integrates to 1 over the sphere
    dirs, w = _sphere_grid(720, 720, upper=False)
    dirs = dirs[:, [0, 2, 1]]  # polar axis y
    p = np.array([oracle.background_evaluate(bs, x)[1] for x in dirs.astype(np.float32)])
    assert float(np.sum(p * w)) == pytest.approx(1.0, rel=5e-3)


def test_colored_background_closed_form(oracle):
    box = np.array([-1, 0, -2, 3, 2, 1], np.float32)
    bg = hip.background_record(hip.BACKGROUND_COLORED, color=(0.5, 1.0, 2.0), scale=(3.0, 2.0, 0.5))
    flux, sat = hip.background_build(bg, None, box)
    dx, dy, dz = 4.0, 2.0, 3.0
    expect = 4 * np.pi * 2 * np.pi * (dx * dy + dx * dz + dy * dz) * np.array([0.5, 1.0, 2.0])
    np.testing.assert_allclose(flux, expect, rtol=1e-6)
    tx = oracle.Textures()
    bs = oracle.background_scene(bg[0], tx, box, flux=flux)
    d, pdf, rad = oracle.background_sample(bs, 0.3, 0.8)
    assert pdf == pytest.approx(1 / (4 * np.pi)) and np.allclose(rad, [1.5, 2.0, 1.0])
    assert abs(np.linalg.norm(d) - 1) < 1e-6
    rad_e, pdf_e = oracle.background_evaluate(bs, d)
    assert np.allclose(rad_e, rad) and pdf_e == pytest.approx(1 / (4 * np.pi))


def test_envmap_scale_per_channel(oracle):
    """The background scale is a spectrum (Background::set_scale,
    background.hpp:119-123): radiance scales per channel, and the env-map flux
    is scaled twice (once per texel in compute_envmap_flux, once more in
    acquire_const, background.cpp:46-54,78-122), so flux(s) = flux(1) * s^2."""
    env = _env()
    box = np.array([-1, 0, -1, 1, 2, 1], np.float32)
    s = np.array([1.5, 0.5, 2.0], np.float32)
    bg1 = hip.background_record(hip.BACKGROUND_ENVMAP, scale=1.0, envmap_texture=0)
    bgs = hip.background_record(hip.BACKGROUND_ENVMAP, scale=s, envmap_texture=0)
    f1, sat1 = hip.background_build(bg1, env, box)
    fs, sats = hip.background_build(bgs, env, box)
    np.testing.assert_allclose(fs, np.asarray(f1, np.float64) * s.astype(np.float64) ** 2, rtol=2e-6)
    assert np.array_equal(sat1, sats)  # the sampling table ignores the scale
    tx = oracle.Textures([(env, hip.SAMPLING_NEAREST)])
    b1 = oracle.background_scene(bg1[0], tx, box, sat=sat1, flux=f1)
    bs = oracle.background_scene(bgs[0], tx, box, sat=sats, flux=fs)
    rng = np.random.default_rng(5)
    for _ in range(50):
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        r1, p1 = oracle.background_evaluate(b1, d)
        rs, ps = oracle.background_evaluate(bs, d)
        np.testing.assert_allclose(rs, r1 * s, rtol=1e-6)
        assert ps == p1
