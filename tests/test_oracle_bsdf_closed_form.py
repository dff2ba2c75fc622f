"""The oracle's evalDisneyBsdf (materials/disney_bsdf.cuh:209-321) against the
published formulas it is made of, written here from the publications in
FP64 (no reference values exist: disney_bsdf.cuh needs CUDA):

* diffuse: Burley 2012 ("Physically-Based Shading at Disney", sec. 5.3):
  base/pi (1 + (F_D90 - 1)(1 - cos_l)^5)(1 + (F_D90 - 1)(1 - cos_v)^5),
  F_D90 = 0.5 + 2 roughness cos_d^2, cos_d = l.h;
* sheen: the same notes: sheen * lerp(1, tint, sheenTint) * (1 - cos_d)^5,
  tint = base / luminance(base) (Rec. 709 weights), times the
  KHR_materials_sheen colour;
* specular and clearcoat: GGX / Trowbridge-Reitz D (Walter et al. 2007),
  height-correlated Smith G2 = 1 / (1 + Lambda(v) + Lambda(l)) with
  Lambda = (sqrt(1 + alpha^2 tan^2) - 1) / 2 (Heitz 2014), Schlick's
  Fresnel, D G F / (4 cos_l cos_v); alpha = roughness^2; F0 = the Fresnel
  reflectance at normal incidence ((ior - 1)/(ior + 1))^2 for dielectrics
  (KHR_materials_specular factor and colour), the base colour for metals;
* pdf: the visible-normal density D_v(h) = G1(v) max(0, v.h) D(h) / cos_v
  (Heitz 2018) over 4 v.h for the specular lobe, cos / pi for diffuse and
  sheen, D cos_h / (4 v.h) for clearcoat, mixed with the reference's lobe
  selection weights (disney_bsdf.cuh:184-205);
* transmission: Walter et al. 2007's rough-dielectric BTDF with the exact
  unpolarised Fresnel reflectance, and its refracted density (see the test
  for the two ways the reference departs from it).

eval returns f times cos_l.  A wrong alpha mapping, a dropped 1/4 or pi, a
Smith term of the wrong form or a Fresnel on the wrong angle fails these.
Anisotropic materials: D is the anisotropic GGX with Burley's aspect
mapping of (roughness, anisotropic) to (alpha_x, alpha_y); the reference
evaluates G with alpha = sqrt(ax ay), not the anisotropic Lambda of the
literature.  The
reference adds 1e-7 to several denominators; inputs keep the relative
effect below 1e-5."""
import numpy as np
import pytest

import oracle
from test_oracle_furnace import _material

LUM = np.array([0.2126, 0.7152, 0.0722])


def _ggx_d(cos_h, a):
    return a * a / (np.pi * (cos_h * cos_h * (a * a - 1.0) + 1.0) ** 2)


def _smith_lambda(c, a):
    tan2 = (1.0 - c * c) / (c * c)
    return (np.sqrt(1.0 + a * a * tan2) - 1.0) / 2.0


def _schlick(f0, c):
    return f0 + (1.0 - f0) * (1.0 - c) ** 5


def _textbook(m, wo, wi, n):
    base = np.array([m.base_color.x, m.base_color.y, m.base_color.z], np.float64)
    cl, cv = wi @ n, wo @ n
    h = (wo + wi) / np.linalg.norm(wo + wi)
    ch, cd = h @ n, h @ wi
    rough = np.clip(m.roughness, 0.001, 1.0)
    met = m.metallic
    tint = base / (LUM @ base)
    # diffuse and sheen
    fd90 = 0.5 + 2.0 * rough * cd * cd
    diffuse = base / np.pi * (1 + (fd90 - 1) * (1 - cl) ** 5) * (1 + (fd90 - 1) * (1 - cv) ** 5)
    sheen_col = ((1 - m.sheen_tint) + m.sheen_tint * tint) * np.array([m.sheen_color_r, m.sheen_color_g, m.sheen_color_b])
    sheen = m.sheen * sheen_col * (1 - cd) ** 5
    # specular
    a = max(rough * rough, 0.001)
    g2 = 1.0 / (1.0 + _smith_lambda(cl, a) + _smith_lambda(cv, a))
    r0 = ((m.ior - 1.0) / (m.ior + 1.0)) ** 2
    f0_d = r0 * ((1 - m.specular_tint) + m.specular_tint * tint) * m.specular_factor * np.array(
        [m.specular_color_r, m.specular_color_g, m.specular_color_b])
    f0 = (1 - met) * f0_d + met * base
    spec = _ggx_d(ch, a) * g2 * _schlick(f0, cd) / (4 * cl * cv)
    # clearcoat
    ac = max(np.clip(m.clearcoat_roughness, 0.001, 1.0) ** 2, 0.001)
    gc = 1.0 / (1.0 + _smith_lambda(cl, ac) + _smith_lambda(cv, ac))
    cc = 0.25 * m.clearcoat * _ggx_d(ch, ac) * gc * _schlick(0.04, cd) / (4 * cl * cv)
    f = ((1 - met) * (1 - m.spec_trans) * diffuse + (1 - met) * sheen * (m.sheen > 0) + spec + cc) * cl
    # pdf: lobe densities mixed by the reference's selection weights
    g1 = 1.0 / (1.0 + _smith_lambda(cv, a))
    vndf = g1 * max(0.0, wo @ h) * _ggx_d(ch, a) / cv
    w = np.array([(1 - met) * (1 - m.spec_trans), 1.0, 0.25 * m.clearcoat, (1 - met) * m.spec_trans,
                  (1 - met) * m.sheen * 0.1])
    w /= w.sum()
    pdf = (w[0] * cl / np.pi * (w[0] > 0) + w[1] * vndf / (4 * cd) + w[2] * _ggx_d(ch, ac) * ch / (4 * cd) * (m.clearcoat > 0)
           + w[4] * cl / np.pi * (m.sheen > 0 and met < 1))
    return f, pdf


CASES = {
    "rough-metal": dict(metallic=1.0, roughness=0.7),
    "dielectric": dict(metallic=0.0, roughness=0.6, specular_tint=0.4, ior=1.6),
    "dielectric-clearcoat-sheen": dict(metallic=0.0, roughness=0.8, clearcoat=0.7, clearcoat_roughness=0.6,
                                       sheen=0.8, sheen_tint=0.3),
    "half-metal-coated": dict(metallic=0.5, roughness=0.55, clearcoat=1.0, clearcoat_roughness=0.5, sheen=0.5),
}


def _upper(rng, n, min_cos=0.15):
    while True:
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        if v @ n >= min_cos:
            return v


@pytest.mark.parametrize("case", sorted(CASES))
def test_eval_is_the_published_disney_ggx_model(case):
    mat = _material(**CASES[case])
    mat.sheen_color_r, mat.sheen_color_g, mat.sheen_color_b = 0.9, 0.7, 1.0
    mat.specular_color_r, mat.specular_color_g, mat.specular_color_b = 1.0, 0.9, 0.8
    rng = np.random.default_rng(7)
    for n in (np.array([0.0, 0.0, 1.0]), np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8])):
        for _ in range(40):
            wo, wi = _upper(rng, n), _upper(rng, n)
            f, pdf = oracle.bsdf_eval(mat, wo, wi, n, True, fp64=True)
            ef, epdf = _textbook(mat, wo, wi, n)
            assert np.allclose(f, ef, rtol=2e-5, atol=0), (case, wo, wi, f, ef)
            assert pdf == pytest.approx(epdf, rel=2e-5), (case, wo, wi, pdf, epdf)
            # the single-precision evaluation the renderer uses stays within float rounding
            f32, pdf32 = oracle.bsdf_eval(mat, wo, wi, n, True, fp64=False)
            assert np.allclose(f32, ef, rtol=2e-4) and pdf32 == pytest.approx(epdf, rel=2e-4)


def test_eval_at_normal_incidence():
    # wo = wi = n: the diffuse retro-reflection and sheen factors vanish and every
    # microfacet lobe reduces to D(n) F0 / 4 = F0 / (4 pi alpha^2)
    n = np.array([0.0, 0.0, 1.0])
    mat = _material(metallic=1.0, roughness=0.6)
    f, pdf = oracle.bsdf_eval(mat, n, n, n, True, fp64=True)
    a = 0.36
    assert np.allclose(f, np.array([0.8, 0.6, 0.4]) / (4 * np.pi * a * a), rtol=1e-5)
    assert pdf == pytest.approx(1.0 / (4 * np.pi * a * a), rel=1e-5)  # G1(n) = 1, v.h = 1
    mat = _material(metallic=0.0, roughness=0.6, specular_factor=0.0)
    f, _ = oracle.bsdf_eval(mat, n, n, n, True, fp64=True)
    assert np.allclose(f, np.array([0.8, 0.6, 0.4]) / np.pi, rtol=1e-5)  # Lambert: base / pi


def _fresnel_dielectric(ci, eta):
    """unpolarised Fresnel reflectance, relative index eta = n_incident / n_transmitted."""
    s2t = eta * eta * (1.0 - ci * ci)
    if s2t >= 1.0:
        return 1.0
    ct = np.sqrt(1.0 - s2t)
    r_perp = (eta * ci - ct) / (eta * ci + ct)
    r_par = (ci - eta * ct) / (ci + eta * ct)
    return 0.5 * (r_perp ** 2 + r_par ** 2)


@pytest.mark.parametrize("roughness", [0.5, 0.8])
def test_transmission_is_walters_rough_dielectric_btdf(roughness):
    # Walter et al. 2007, "Microfacet models for refraction through rough surfaces", eq. 21
    # with h_t = -(eta wo + wi) / |.|:  f_t = |wo.h||wi.h| / (|wo.n||wi.n|) (1 - F) G D / (eta wo.h + wi.h)^2,
    # and the refracted density D_v(h) |wi.h| / (eta wo.h + wi.h)^2 (eq. 17's Jacobian).  Two properties of
    # the reference stand beside them: transmitted f carries no cosine (the reflection branch multiplies
    # by cos_l, disney_bsdf.cuh:319) and the pdf has one more eta^2 (disney_bsdf.cuh:247-250)
    mat = _material(metallic=0.2, roughness=roughness, spec_trans=0.9, ior=1.5)
    base = np.array([0.8, 0.6, 0.4])
    n = np.array([0.0, 0.0, 1.0])
    eta = 1.0 / 1.5
    a = roughness * roughness
    w = np.array([(1 - 0.2) * (1 - 0.9), 1.0, 0.0, (1 - 0.2) * 0.9, 0.0])
    w_t = w[3] / w.sum()
    rng = np.random.default_rng(5)
    checked = 0
    while checked < 60:
        wo = _upper(rng, n)
        wi = -_upper(rng, n)
        h = -(eta * wo + wi)
        h /= np.linalg.norm(h)
        den = eta * (wo @ h) + wi @ h
        if abs(den) < 0.1:
            continue
        cv, cl = abs(wo @ n), abs(wi @ n)
        D = _ggx_d(abs(h @ n), a)
        G = 1.0 / (1.0 + _smith_lambda(cl, a) + _smith_lambda(cv, a))
        F = _fresnel_dielectric(abs(wo @ h), eta)
        ft = base * (1 - 0.2) * 0.9 * abs(wo @ h) * abs(wi @ h) / (cv * cl) * (1 - F) * G * D / den ** 2
        g1 = 1.0 / (1.0 + _smith_lambda(cv, a))
        pdf = w_t * g1 * abs(wo @ h) * D / cv * abs(wi @ h) / den ** 2 * eta ** 2
        f, p = oracle.bsdf_eval(mat, wo, wi, n, True, fp64=True)
        assert np.allclose(f, ft, rtol=2e-5), (wo, wi, f, ft)
        assert p == pytest.approx(pdf, rel=2e-5), (wo, wi, p, pdf)
        checked += 1


@pytest.mark.parametrize("aniso", [0.3, 0.8])
def test_anisotropic_alpha_is_burleys_aspect_mapping(aniso):
    # Burley 2012 (addendum B): aspect = sqrt(1 - 0.9 anisotropic), alpha_x =
    # r^2 / aspect, alpha_y = r^2 aspect; D is anisotropic GGX in the frame
    # math/frame.h builds around N = +z (tangent +y, bitangent -x); the
    # reference's G and VNDF G1 use the isotropic alpha sqrt(alpha_x alpha_y).
    # A metal isolates the specular lobe (its selection weight is 1).
    r = 0.6
    mat = _material(metallic=1.0, roughness=r, anisotropic=aniso)
    base = np.array([0.8, 0.6, 0.4])
    n = np.array([0.0, 0.0, 1.0])
    t, b = np.array([0.0, 1.0, 0.0]), np.array([-1.0, 0.0, 0.0])
    aspect = np.sqrt(1.0 - 0.9 * aniso)
    ax, ay = max(r * r / aspect, 0.001), max(r * r * aspect, 0.001)
    ag = np.sqrt(ax * ay)
    rng = np.random.default_rng(13)
    for _ in range(60):
        wo, wi = _upper(rng, n), _upper(rng, n)
        h = (wo + wi) / np.linalg.norm(wo + wi)
        hx, hy, hz = h @ t, h @ b, h @ n
        D = 1.0 / (np.pi * ax * ay * ((hx / ax) ** 2 + (hy / ay) ** 2 + hz ** 2) ** 2)
        cl, cv = wi @ n, wo @ n
        G = 1.0 / (1.0 + _smith_lambda(cl, ag) + _smith_lambda(cv, ag))
        F = _schlick(base, h @ wi)
        ef = D * G * F / (4 * cl * cv) * cl
        g1 = 1.0 / (1.0 + _smith_lambda(cv, ag))
        epdf = g1 * max(0.0, wo @ h) * D / cv / (4 * (h @ wi))
        f, pdf = oracle.bsdf_eval(mat, wo, wi, n, True, fp64=True)
        assert np.allclose(f, ef, rtol=2e-5), (wo, wi, f, ef)
        assert pdf == pytest.approx(epdf, rel=2e-5)
