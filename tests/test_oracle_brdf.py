"""Oracle Principled BSDF.

Part 1 re-expresses PathTracerLib/tests/PathTracerPrincipledBrdfTests.cpp:80-387 (same materials,
same wo = normalize(0.2, 0.1, +-1), n = +z, same u values, same tolerances).
Part 2 pins the lobes with closed forms and integrals that a dropped term, a wrong sign or a
transposed operand would break.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

WO_UP = O.Vec3(0.2, 0.1, 1.0)
WO_DOWN = O.Vec3(0.2, 0.1, -1.0)


@pytest.fixture(scope="module")
def L():
    return O.lib()


def mk(L, mat, inside=False, eta=1.0, flags=0):
    wo = L.orc_vec_normalize(WO_DOWN if inside else WO_UP)
    return O.ctx(mat, (0.0, 0.0, 1.0), wo, eta, flags)


def M(r, g, b, roughness=0.5, metallic=0.0, transmission=0.0, thin=0.0, ior=1.5, subsurface=0.0, emission=0.0,
      diffuseRoughness=-1.0, sr=0.0, sg=0.0, sb=0.0, specular=1.0):
    return O.material(r, g, b, roughness, metallic, transmission, thin, ior, subsurface, emission, diffuseRoughness,
                      sr, sg, sb, specular)


def dot(a, b):
    return a.x * b.x + a.y * b.y + a.z * b.z


def throughput(L, c, s):
    f = L.orc_brdf_eval(c, s.direction)
    return L.orc_apply_throughput(O.Vec3(1, 1, 1), c, s, f)


# ---- Part 1: the reference's property tests --------------------------------------------------

def test_transmissive_sample_throughput_near_one(L):           # :80-93
    c = mk(L, M(1, 1, 1, 0, 0, 1, 0))
    s = L.orc_brdf_sample(c, 0.6, 0.5)
    assert s.valid and s.transmitted
    assert abs(throughput(L, c, s).x - 1.0) <= 0.15


def test_transmissive_reflected_throughput_near_one(L):         # :95-108
    c = mk(L, M(1, 1, 1, 0, 0, 1, 0))
    s = L.orc_brdf_sample(c, 0.99, 0.5)
    assert s.valid and not s.transmitted
    assert abs(throughput(L, c, s).x - 1.0) <= 0.15


def test_thin_transmissive_sample(L):                            # :110-120
    c = mk(L, M(0.2, 0.8, 0.1, 0.5, 0, 1, 1))
    s = L.orc_brdf_sample(c, 0.6, 0.5)
    assert s.valid and s.transmitted
    wo = c.wo
    assert abs(dot(s.direction, O.Vec3(-wo.x, -wo.y, -wo.z)) - 1.0) <= 0.01


def test_subsurface_sample(L):                                   # :122-133
    c = mk(L, M(0.9, 0.85, 0.7, 0.85, 0, 0, 0, 1.5, 0.8, 0, 0.85, 0.05, 0.04, 0.02))
    s = L.orc_brdf_sample(c, 0.95, 0.25)
    assert s.valid and s.direction.z > 0.0 and s.subsurfaceScatter


def test_metallic_sample(L):                                     # :135-144
    c = mk(L, M(0.9, 0.9, 0.9, 0.15, 1.0))
    s = L.orc_brdf_sample(c, 0.25, 0.75)
    assert s.valid and s.direction.z > 0.0


def test_diffuse_throughput(L):                                  # :146-163
    c = mk(L, M(0.72, 0.68, 0.58, 0.85))
    for u1 in np.arange(0.1, 0.9, 0.2, dtype=np.float32):
        for u2 in np.arange(0.1, 0.9, 0.2, dtype=np.float32):
            s = L.orc_brdf_sample(c, float(u1), float(u2))
            if not s.valid:
                continue
            f = L.orc_brdf_eval(c, s.direction)
            scale = max(0.0, s.direction.z) * f.x / s.pdf
            assert 0.1 < scale < 1.5


def test_skips_environment_nee(L):                              # :165-171
    assert L.orc_skips_env_nee(M(1, 1, 1, 0, 0, 1))
    assert not L.orc_skips_env_nee(M(0.72, 0.68, 0.58))


def test_lobe_weights_from_parameters(L):                        # :173-180
    w = L.orc_lobe_weights(mk(L, M(1, 1, 1, 0, 0, 1)))
    assert w.transmit >= 0.99 and w.specular < 0.01


def test_refracted_direction_nonzero(L):                         # :182-197
    c = mk(L, M(1, 1, 1, 0, 0, 1, 0))
    d = O.Vec3()
    a, b, nx = (O.C.c_float() for _ in range(3))
    assert L.orc_refract_direction(c, O.C.byref(d), O.C.byref(a), O.C.byref(b), O.C.byref(nx))
    assert L.orc_brdf_eval(c, d).x > 0.0


def test_sample_transmit_forced_refract(L):                      # :199-210
    c = mk(L, M(1, 1, 1, 0, 0, 1, 0, 1.5))
    s = L.orc_brdf_sample_transmit(c, 0.0, 0.5)
    assert s.valid and s.transmitted and abs(s.nextMediumEta - 1.5) <= 0.01 and s.direction.z < 0.0


def test_sample_transmit_forced_reflect(L):                      # :212-223
    c = mk(L, M(1, 1, 1, 0, 0, 1, 0, 1.5))
    s = L.orc_brdf_sample_transmit(c, 1.0, 0.5)
    assert s.valid and not s.transmitted and abs(s.nextMediumEta - 1.0) <= 0.01 and s.direction.z > 0.0


def test_debug_disable_refract(L):                               # :225-236
    c = mk(L, M(1, 1, 1, 0, 0, 1, 0, 1.5), flags=1 << 1)
    s = L.orc_brdf_sample_transmit(c, 0.0, 0.5)
    assert s.valid and not s.transmitted and s.direction.z > 0.0


def test_debug_disable_reflect(L):                               # :238-249
    c = mk(L, M(1, 1, 1, 0, 0, 1, 0, 1.5), flags=1 << 2)
    s = L.orc_brdf_sample_transmit(c, 1.0, 0.5)
    assert s.valid and s.transmitted and abs(s.nextMediumEta - 1.5) <= 0.01


def test_debug_force_transmit_lobe_only(L):                      # :251-264
    c = mk(L, M(1, 1, 1, 0, 0, 1, 0, 1.5), flags=1 << 0)
    s = L.orc_brdf_sample(c, 0.5, 0.5)
    assert s.valid and (s.transmitted or s.direction.z > 0.0)


@pytest.mark.parametrize("ior", [1.5, 1.33])
def test_glass_medium_eta_enter_exit(L, ior):                    # :266-281, :337-357
    m = M(1, 1, 1, 0, 0, 1, 0, ior)
    enter = L.orc_brdf_sample_transmit(mk(L, m, eta=1.0), 0.0, 0.5)
    assert enter.transmitted and abs(enter.nextMediumEta - ior) <= 0.01
    exit_ = L.orc_brdf_sample_transmit(mk(L, m, inside=True, eta=enter.nextMediumEta), 0.0, 0.5)
    assert exit_.transmitted and abs(exit_.nextMediumEta - 1.0) <= 0.01
    again = L.orc_brdf_sample_transmit(mk(L, m, eta=exit_.nextMediumEta), 0.0, 0.5)
    assert again.transmitted and abs(again.nextMediumEta - ior) <= 0.01


def test_integrator_glass_state(L):                              # :283-302
    c = mk(L, M(1, 1, 1, 0, 0, 1, 0, 1.5))
    r = L.orc_brdf_sample_transmit(c, 0.0, 0.5)
    assert r.transmitted and abs(r.nextMediumEta - 1.5) <= 0.01 and r.direction.z < 0.0
    f = L.orc_brdf_sample_transmit(c, 1.0, 0.5)
    assert not f.transmitted and abs(f.nextMediumEta - c.etaMedium) <= 0.01 and f.direction.z > 0.0


def test_oren_nayar_zero_roughness_is_lambert(L):                # :304-311
    c = mk(L, M(0.8, 0.8, 0.8, 0.5, 0, 0, 0, 1.5, 0, 0, 0.0))
    assert abs(L.orc_oren_nayar_ctx(c, L.orc_vec_normalize(O.Vec3(0.3, 0.2, 1.0))) - 1.0) <= 0.05


def test_oren_nayar_high_roughness_broadens(L):                  # :313-323
    wi = L.orc_vec_normalize(O.Vec3(0.5, 0.1, 1.0))
    smooth = L.orc_oren_nayar_ctx(mk(L, M(0.8, 0.8, 0.8, 0.1, 0, 0, 0, 1.5, 0, 0, 0.0)), wi)
    rough = L.orc_oren_nayar_ctx(mk(L, M(0.8, 0.8, 0.8, 0.1, 0, 0, 0, 1.5, 0, 0, 0.95)), wi)
    assert abs(smooth - 1.0) <= 0.05 and abs(rough - smooth) > 0.05


def test_metallic_no_diffuse_leak(L):                            # :325-333
    w = L.orc_lobe_weights(mk(L, M(0.9, 0.7, 0.2, 0.2, 1.0)))
    assert w.diffuse < 0.01 and w.transmit < 0.01 and w.subsurface < 0.01


def test_metallic_energy_compensation(L):                        # :335-345
    c = mk(L, M(0.9, 0.9, 0.9, 0.85, 1.0))
    s = L.orc_brdf_sample(c, 0.3, 0.6)
    assert s.valid and L.orc_brdf_eval(c, s.direction).x > 0.01


def test_thin_tinted_throughput(L):                              # :359-369
    c = mk(L, M(0.2, 0.8, 0.1, 0.5, 0, 1, 1))
    s = L.orc_brdf_sample(c, 0.1, 0.5)
    assert s.valid and s.transmitted
    t = throughput(L, c, s)
    assert t.y > t.x


def test_rough_glass_samples_non_delta(L):                       # :371-380
    s = L.orc_brdf_sample_transmit(mk(L, M(1, 1, 1, 0.35, 0, 1, 0, 1.5)), 0.05, 0.37)
    assert s.valid and s.transmitted


# ---- Part 2: closed forms and integrals ---------------------------------------------------------

@pytest.mark.parametrize("eta", [1.33, 1.5, 2.4])
def test_fresnel_normal_incidence(L, eta):
    want = ((eta - 1) / (eta + 1)) ** 2
    assert abs(L.orc_dielectric_fresnel(1.0, 1.0, eta) - want) < 2e-7
    # total internal reflection past the critical angle from inside
    crit = math.sqrt(1 - 1 / eta**2)
    assert L.orc_dielectric_fresnel(-(crit * 0.5), 1.0, eta) == 1.0


def test_schlick_endpoints(L):
    f0 = O.Vec3(0.04, 0.5, 0.9)
    a = L.orc_schlick_f(1.0, f0)
    b = L.orc_schlick_f(0.0, f0)
    assert (a.x, a.y, a.z) == (np.float32(0.04), np.float32(0.5), np.float32(0.9))
    assert (b.x, b.y, b.z) == (1.0, 1.0, 1.0)


@pytest.mark.parametrize("alpha", [0.0016, 0.04, 0.25, 0.81])
def test_ggx_normalisation(L, alpha):
    """integral over the hemisphere of D(h) cos(theta_h) = 1."""
    # geometric grid in 1 - cos(theta_h): resolves the narrow peak of small alpha
    s = np.concatenate([[0.0], np.geomspace(1e-12, 1.0, 20000)])
    mu = 1.0 - s
    d = np.array([L.orc_ggx_d(float(np.float32(m)), alpha) for m in mu], dtype=np.float64)
    integral = 2 * math.pi * np.trapezoid(d * mu, s)
    assert abs(integral - 1.0) < 2e-3


def test_smith_g1_at_normal(L):
    for a in (0.01, 0.3, 0.9):
        assert abs(L.orc_smith_g1(1.0, a) - 1.0) < 1e-6
        assert L.orc_smith_g1(0.2, a) < 1.0


def test_energy_compensation_identity(L):
    assert L.orc_energy_compensation(0.6, 1.0) == 1.0
    assert L.orc_energy_compensation(0.9, 0.2) > 1.0


def test_oren_nayar_sigma_zero(L):
    assert L.orc_oren_nayar_factor(0.0, 0.3, 0.8, 0.1) == 1.0


def test_cosine_sample_closed_form(L):
    for u1, u2 in [(0.1, 0.2), (0.7, 0.9), (0.5, 0.5)]:
        v = L.orc_sample_cosine_local(u1, u2)
        r = math.sqrt(u2)
        phi = 2 * math.pi * u1
        assert abs(v.x - r * math.cos(phi)) < 1e-6 and abs(v.y - r * math.sin(phi)) < 1e-6
        assert abs(v.z - math.sqrt(1 - u2)) < 1e-6


def hemisphere_grid(n_theta=256, n_phi=512, lower=False):
    ct = (np.arange(n_theta) + 0.5) / n_theta
    if lower:
        ct = -ct
    phi = (np.arange(n_phi) + 0.5) / n_phi * 2 * math.pi
    C, P = np.meshgrid(ct, phi, indexing="ij")
    S = np.sqrt(1 - C**2)
    dirs = np.stack([S * np.cos(P), S * np.sin(P), C], -1).reshape(-1, 3)
    dw = 2 * math.pi / (n_theta * n_phi)
    return dirs.astype(np.float32), dw


@pytest.mark.parametrize("mat", [dict(), dict(roughness=0.3, metallic=1.0), dict(roughness=0.9),
                                 dict(roughness=0.6, diffuseRoughness=0.9, specular=0.5)])
def test_pdf_integrates_to_one(L, mat):
    m = M(0.4, 0.3, 0.2, **mat)
    c = mk(L, m)
    dirs, dw = hemisphere_grid()
    p = np.array([L.orc_brdf_pdf(c, O.v3(d)) for d in dirs])
    total = p.sum() * dw
    w = L.orc_lobe_weights(c)
    # the diffuse lobe's pdf integrates to its weight; the GGX reflection pdf integrates to its
    # weight minus the mass of half-vectors that reflect below the horizon
    assert w.diffuse - 1e-3 <= total <= 1.0 + 1e-3


def test_sampling_matches_pdf(L):
    """Sampled directions are distributed as brdf_pdf: the importance-sampled estimate of
    integral of pdf(w) * g(w) dw / pdf = E[g] equals the quadrature of pdf * g."""
    c = mk(L, M(0.4, 0.3, 0.2, 0.6))
    rng = np.random.default_rng(11)
    u = rng.random((40000, 2)).astype(np.float32)
    g = lambda d: d[2] ** 2  # noqa: E731
    vals = []
    for u1, u2 in u:
        s = L.orc_brdf_sample(c, float(u1), float(u2))
        vals.append(g((s.direction.x, s.direction.y, s.direction.z)) if s.valid else 0.0)
    mc = np.mean(vals)
    dirs, dw = hemisphere_grid()
    p = np.array([L.orc_brdf_pdf(c, O.v3(d)) for d in dirs])
    quad = np.sum(p * dirs[:, 2] ** 2) * dw
    assert abs(mc - quad) < 4 * np.std(vals) / math.sqrt(len(vals)) + 1e-3


def test_diffuse_albedo_lambert_limit(L):
    """With zero diffuse roughness and no specular lobe left (dielectric weights), the diffuse
    lobe integrates to weight * base colour (Lambert: integral of base/pi cos = base)."""
    m = M(0.4, 0.3, 0.2, roughness=0.5, diffuseRoughness=0.0, specular=0.0)
    c = mk(L, m)
    w = L.orc_lobe_weights(c)
    dirs, dw = hemisphere_grid()
    f = np.array([L.orc_brdf_eval(c, O.v3(d)).x for d in dirs])
    spec_free = f - 0.0
    # specular=0 removes F0, leaving only the Schlick grazing term; bound it from both sides
    alb = np.sum(spec_free * dirs[:, 2]) * dw
    assert w.diffuse * 0.4 <= alb <= w.diffuse * 0.4 + 0.05


def test_mis_balance(L):
    assert L.orc_mis_balance(1.0, 3.0) == 0.25
    assert L.orc_mis_balance(0.0, 1.0) == 0.0
    assert L.orc_mis_balance(2.0, -2.0) == 0.0
    ref = O.ref_policy()
    if ref is not None:
        rng = np.random.default_rng(0)
        for a, b in rng.normal(size=(300, 2)).astype(np.float32):
            assert L.orc_mis_balance(float(a), float(b)) == ref.ref_mis_balance(float(a), float(b))


def test_throughput_is_f_cos_over_pdf(L):
    c = mk(L, M(0.4, 0.3, 0.2, 0.6))
    s = L.orc_brdf_sample(c, 0.2, 0.7)
    f = L.orc_brdf_eval(c, s.direction)
    t = L.orc_apply_throughput(O.Vec3(0.5, 1.0, 2.0), c, s, f)
    k = np.float32(np.float32(max(0.0, s.direction.z)) / np.float32(s.pdf))
    assert t.x == np.float32(np.float32(0.5) * np.float32(np.float32(f.x) * k))
