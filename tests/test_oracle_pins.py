"""Pins the oracle's microfacet BSDFs (conductor, rough dielectric,
principled) and its point lights against closed forms taken from the
literature, written here in float64 numpy from the published formulas and
not from the oracle's code.  CPU only.

Textbook forms used:
- Trowbridge-Reitz (GGX) D, anisotropic: D(h) = 1 / (pi ax ay ((hx/ax)^2 +
  (hy/ay)^2 + hz^2)^2) (Heitz 2014, "Understanding the masking-shadowing
  function", eq. 85 rearranged);
- Smith Lambda(w) = (sqrt(1 + (ax^2 wx^2 + ay^2 wy^2) / wz^2) - 1) / 2,
  G1 = 1 / (1 + Lambda), height-correlated G2 = 1 / (1 + Lambda(wo) +
  Lambda(wi)) (Heitz 2014 eqs. 86, 99);
- visible-normal pdf of a reflected direction G1(wo) D(h) / (4 wo.z) (Heitz
  2018, "Sampling the GGX distribution of visible normals", eq. 3 with the
  reflection Jacobian 1 / (4 wo.h));
- conductor Fresnel from the complex index eta + i k (Fresnel equations with
  complex Snell's law, unpolarised average of |r_s|^2 and |r_p|^2);
- dielectric Fresnel from Snell's law (unpolarised, real);
- Walter et al. 2007 ("Microfacet models for refraction through rough
  surfaces", eqs. 16-21) for the rough dielectric BTDF and its pdf.

The reference's principled diffuse lobe is not Burley's (its retro-reflection
term rr (F_i + F_o F_i F_o (rr - 1)) and its Fresnel weights 1 + (fd90 - 1)
(1 - cos)^5 differ from the 2012 notes, principled.glsl:50-74), so no
textbook form exists for it: it is isolated by subtracting the textbook
specular lobe and compared with the reference's formula at direction pairs
where a swap of wi and wo, a dropped factor or a wrong exponent changes the
value.
"""
import math

import numpy as np
import pytest

import oracle
from lumen_amd import _abi

PI = math.pi


def _mat(**kw):
    m = _abi.lt_material()
    m.texture_id = -1
    for k, v in kw.items():
        if isinstance(v, (tuple, list)):
            getattr(m, k)[:] = list(v)
        else:
            setattr(m, k, v)
    return m


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


N = np.array([0.0, 0.0, 1.0])


# ------------------------------------------------------------ textbook ----
def D_tr(h, ax, ay):
    return 1.0 / (PI * ax * ay * ((h[0] / ax) ** 2 + (h[1] / ay) ** 2 + h[2] ** 2) ** 2)


def lam(w, ax, ay):
    return 0.5 * (math.sqrt(1.0 + (ax * ax * w[0] ** 2 + ay * ay * w[1] ** 2) / w[2] ** 2) - 1.0)


def G1(w, ax, ay):
    return 1.0 / (1.0 + lam(w, ax, ay))


def G2(wo, wi, ax, ay):
    return 1.0 / (1.0 + lam(wo, ax, ay) + lam(wi, ax, ay))


def fresnel_conductor_complex(c, eta, k):
    n = complex(eta, k)
    s2 = 1.0 - c * c
    ct = np.sqrt(1.0 - s2 / (n * n))
    rs = (c - n * ct) / (c + n * ct)
    rp = (n * c - ct) / (n * c + ct)
    return 0.5 * (abs(rs) ** 2 + abs(rp) ** 2)


def fresnel_dielectric_snell(c, eta):
    """Unpolarised reflectance at a boundary from index 1 into index eta
    (eta < 1: from inside), c = cos of the incident angle."""
    s2t = (1.0 - c * c) / (eta * eta)
    if s2t >= 1.0:
        return 1.0
    ct = math.sqrt(1.0 - s2t)
    rs = (c - eta * ct) / (c + eta * ct)
    rp = (eta * c - ct) / (eta * c + ct)
    return 0.5 * (rs * rs + rp * rp)


def schlick(f0, c):
    return f0 + (1.0 - f0) * (1.0 - c) ** 5


# direction pairs in the upper hemisphere (n = +z), deliberately asymmetric
PAIRS = [
    (_unit([0.3, -0.2, 0.9]), _unit([-0.5, 0.4, 0.6])),
    (_unit([0.7, 0.1, 0.3]), _unit([-0.2, -0.6, 0.75])),
    (_unit([0.05, 0.02, 1.0]), _unit([0.6, 0.3, 0.4])),
    (_unit([-0.8, 0.3, 0.25]), _unit([0.7, -0.1, 0.5])),
]


def _eval(m, wo, wi, mode=1, side=True):
    f, pdf, pdf_rev = oracle.eval_bsdf(m, N, wo, wi, mode=mode, side=side)
    return f.astype(np.float64), pdf, pdf_rev


# ----------------------------------------------------------- conductor ----
@pytest.mark.parametrize("c", [1.0, 0.8, 0.5, 0.2, 0.05])
@pytest.mark.parametrize("eta,k", [(0.2, 3.9), (1.1, 2.4), (0.14, 4.2), (2.9, 0.3)])
def test_fresnel_conductor_equals_complex_fresnel(c, eta, k):
    got = oracle.lib().ora_fresnel_conductor(c, eta, k)
    assert got == pytest.approx(fresnel_conductor_complex(c, eta, k), rel=2e-5, abs=1e-6)


@pytest.mark.parametrize("rough", [0.2, 0.45, 0.9])
def test_rough_conductor_is_ggx_times_fresnel(rough):
    eta, k = (0.2, 0.9, 1.1), (3.9, 2.4, 2.2)
    m = _mat(bsdf_type=16, bsdf_props=8 | 4, roughness=rough, albedo=eta, k=k)
    a = rough * rough
    for wo, wi in PAIRS:
        h = _unit(wo + wi)
        F = np.array([fresnel_conductor_complex(float(wo @ h), e, kk) for e, kk in zip(eta, k)])
        want = D_tr(h, a, a) * F * G2(wo, wi, a, a) / (4 * wo[2] * wi[2])
        f, pdf, pdf_rev = _eval(m, wo, wi)
        assert f == pytest.approx(want, rel=2e-4)
        assert pdf == pytest.approx(G1(wo, a, a) * D_tr(h, a, a) / (4 * wo[2]), rel=2e-4)
        assert pdf_rev == pytest.approx(G1(wi, a, a) * D_tr(h, a, a) * (wi @ h) / wi[2] / (4 * (wo @ h)), rel=2e-4)
        assert oracle.bsdf_pdf(m, N, wo, wi) == pytest.approx(pdf, rel=1e-6)


# ---------------------------------------------------- rough dielectric ----
IOR = 1.5


def _walter_t(wo, wi, a, eta, F):
    """Walter 2007 BTDF for radiance leaving along wo (index 1) that arrived
    along wi (index eta): h = -(wo + eta wi) normalised to +z, f = |wi.h||wo.h|
    / (|wi.n||wo.n|) (1 - F) D G / (wo.h + eta wi.h)^2."""
    h = -(wo + eta * wi)
    h = h / np.linalg.norm(h)
    if h[2] < 0:
        h = -h
    den = (wo @ h + eta * (wi @ h)) ** 2
    f = abs(wi @ h) * abs(wo @ h) / (abs(wi[2]) * abs(wo[2])) * (1 - F) * D_tr(h, a, a) * G2(wo, wi, a, a) / den
    # pdf of the sampler: visible normals of wo, transmission Jacobian eta^2 |wi.h| / den
    pdf = G1(wo, a, a) * D_tr(h, a, a) * max(0.0, wo @ h) / wo[2] * (1 - F) * eta * eta * abs(wi @ h) / den
    return h, f, pdf


@pytest.mark.parametrize("rough", [0.3, 0.6])
def test_rough_dielectric_reflection_is_walter(rough):
    m = _mat(bsdf_type=8, bsdf_props=8 | 16 | 4, roughness=rough, ior=IOR, albedo=(1, 1, 1))
    a = rough * rough
    for wo, wi in PAIRS:
        h = _unit(wo + wi)
        F = fresnel_dielectric_snell(float(wo @ h), IOR)
        f, pdf, _ = _eval(m, wo, wi)
        assert f == pytest.approx([F * D_tr(h, a, a) * G2(wo, wi, a, a) / (4 * wo[2] * wi[2])] * 3, rel=3e-4)
        # reflection picked with probability F out of F + (1 - F)
        assert pdf == pytest.approx(G1(wo, a, a) * D_tr(h, a, a) / (4 * wo[2]) * F, rel=3e-4)


@pytest.mark.parametrize("rough", [0.3, 0.6])
@pytest.mark.parametrize("mode", [0, 1])
def test_rough_dielectric_sampled_transmission_is_walter(rough, mode):
    """Transmitted samples of sample_bsdf carry Walter's BTDF (radiance,
    mode 1; eta^2 times that for importance, mode 0) and its pdf, and
    bsdf_pdf agrees.  eval_bsdf returns 0 for transmission: the reference's
    refraction branch returns an unassigned value (dielectric.glsl:171), which
    the oracle restates as 0."""
    m = _mat(bsdf_type=8, bsdf_props=8 | 16 | 4, roughness=rough, ior=IOR, albedo=(1, 1, 1))
    a = rough * rough
    rng = np.random.default_rng(5)
    seen = 0
    for wo in (_unit([0.3, -0.2, 0.9]), _unit([0.7, 0.1, 0.3])):
        for r in rng.random((300, 3)):
            r[0] = 0.5 * r[0] + 0.45  # xi.x also picks the lobe: keep it where transmission wins mostly
            wi, f, pdf, cos = oracle.sample_bsdf(m, N, wo, r, mode=mode)
            if pdf <= 0 or wi[2] >= 0 or not np.any(f):
                continue
            wi = wi.astype(np.float64)
            h = -(wo + IOR * wi)
            h /= np.linalg.norm(h)
            h = h if h[2] > 0 else -h
            F = fresnel_dielectric_snell(float(wo @ h), IOR)
            _, ft, pt = _walter_t(wo, wi, a, IOR, F)
            # Walter's form is the radiance BTDF; importance (mode 0) drops
            # the 1/eta^2 of radiance compression (PBRT-v4 9.7, Veach 1997 5.2)
            scale = 1.0 if mode == 1 else IOR ** 2
            assert f == pytest.approx([ft * scale] * 3, rel=2e-3)
            assert pdf == pytest.approx(pt, rel=2e-3)
            assert oracle.bsdf_pdf(m, N, wo, wi) == pytest.approx(pt, rel=2e-3)
            fe, _, _ = _eval(m, wo, wi, mode=mode)
            assert not np.any(fe)
            seen += 1
    assert seen > 200


# ---------------------------------------------------------- principled ----
def _aniso(rough, an):
    asp = math.sqrt(1 - 0.9 * an)
    return max(0.001, rough * rough / asp), max(0.001, rough * rough * asp)


@pytest.mark.parametrize("rough,an", [(0.3, 0.0), (0.6, 0.0), (0.4, 0.8)])
def test_principled_metal_is_ggx_schlick_heightcorrelated(rough, an):
    """metallic = 1: only the specular lobe, Fresnel = Schlick with F0 = albedo."""
    alb = (0.9, 0.6, 0.3)
    m = _mat(bsdf_type=32, bsdf_props=8 | 4, roughness=rough, metallic=1.0, anisotropy=an, ior=1.5, albedo=alb)
    ax, ay = _aniso(rough, an)
    for wo, wi in PAIRS:
        h = _unit(wo + wi)
        F = np.array([schlick(x, float(wo @ h)) for x in alb])
        want = D_tr(h, ax, ay) * F * G2(wo, wi, ax, ay) / (4 * wo[2] * wi[2])
        f, pdf, _ = _eval(m, wo, wi)
        assert f == pytest.approx(want, rel=3e-4)
        assert pdf == pytest.approx(G1(wo, ax, ay) * D_tr(h, ax, ay) / (4 * wo[2]), rel=3e-4)


def _ref_diffuse(alb, rough, flat, wo, wi):
    """The reference's diffuse lobe (principled.glsl:50-74 with
    disney_fresnel, sampling_commons.glsl:102-110), evaluated in float64."""
    h = _unit(wo + wi)
    fd90 = 0.5 + 2.0 * (wo @ h) ** 2 * rough
    f_wi = 1.0 + (fd90 - 1.0) * (1.0 - wi[2]) ** 5
    f_wo = 1.0 + (fd90 - 1.0) * (1.0 - wo[2]) ** 5
    rr = 2.0 * rough * rough * wi[2] ** 2
    retro = rr * (f_wi + f_wo * f_wi * f_wo * (rr - 1.0))
    diff = (1.0 - 0.5 * f_wi) * (1.0 - 0.5 * f_wo)
    ss = 0.0
    if flat > 0:
        fss90 = 0.5 * rr
        fss = (1 + (fss90 - 1) * f_wi) * (1 + (fss90 - 1) * f_wo)
        ss = 1.25 * (fss * (1.0 / (wi[2] + wo[2]) - 0.5) + 0.5)
    s = (diff + retro) * (1 - flat) + ss * flat
    return np.asarray(alb) * s / PI


@pytest.mark.parametrize("rough,flat", [(0.5, 0.0), (0.8, 0.0), (0.5, 0.6)])
def test_principled_dielectric_specular_plus_reference_diffuse(rough, flat):
    """metallic = 0, no transmission, no clearcoat, no specular tint: the
    specular lobe is GGX x exact dielectric Fresnel x height-correlated G2;
    what remains is the reference's diffuse lobe."""
    alb = (0.8, 0.5, 0.3)
    m = _mat(bsdf_type=32, bsdf_props=8 | 4, roughness=rough, metallic=0.0, ior=IOR, albedo=alb, flatness=flat)
    a = rough * rough
    for wo, wi in PAIRS:
        h = _unit(wo + wi)
        spec = D_tr(h, a, a) * fresnel_dielectric_snell(float(wo @ h), IOR) * G2(wo, wi, a, a) / (4 * wo[2] * wi[2])
        f, pdf, _ = _eval(m, wo, wi)
        assert f - spec == pytest.approx(_ref_diffuse(alb, rough, flat, wo, wi), rel=1e-3, abs=1e-6)
        # lobe probabilities: specular 1, diffuse 1 (brdf weight), normalised to 1/2 each
        want_pdf = 0.5 * G1(wo, a, a) * D_tr(h, a, a) / (4 * wo[2]) + 0.5 * wi[2] / PI
        assert pdf == pytest.approx(want_pdf, rel=3e-4)


def test_principled_diffuse_at_normal_incidence():
    """wo = wi = n with a specular lobe that is a delta (roughness 0.05):
    f = albedo (1/4 + 4 r^4) / pi (Fresnel weights 1, retro rr^2, rr = 2 r^2)."""
    r = 0.05
    m = _mat(bsdf_type=32, bsdf_props=8 | 4, roughness=r, metallic=0.0, ior=IOR, albedo=(1.0, 0.5, 0.25))
    f, _, _ = _eval(m, N, N)
    assert f == pytest.approx(np.array([1.0, 0.5, 0.25]) * (0.25 + 4 * r ** 4) / PI, rel=1e-5)


@pytest.mark.parametrize("gloss", [0.0, 0.7, 1.0])
def test_principled_clearcoat_lobe(gloss):
    """The clearcoat lobe alone (the difference of clearcoat 1 and 0): GGX with
    alpha^2 = mix(0.1, 0.001, gloss), Schlick with F0 = 0.04 at wi.h and the
    separable Smith G1(wo) G1(wi) at alpha = 0.25, times clearcoat / 4."""
    base = dict(bsdf_type=32, bsdf_props=8 | 4, roughness=0.5, metallic=1.0, ior=IOR, albedo=(0.5, 0.5, 0.5),
                clearcoat_gloss=gloss)
    m0, m1 = _mat(**base, clearcoat=0.0), _mat(**base, clearcoat=1.0)
    a = math.sqrt(0.1 + (0.001 - 0.1) * gloss)
    for wo, wi in PAIRS:
        h = _unit(wo + wi)
        want = 0.25 * D_tr(h, a, a) * schlick(0.04, float(wi @ h)) * G1(wo, 0.25, 0.25) * G1(wi, 0.25, 0.25)
        f0, _, _ = _eval(m0, wo, wi)
        f1, pdf1, _ = _eval(m1, wo, wi)
        assert f1 - f0 == pytest.approx([want] * 3, rel=1e-3, abs=1e-7)
        # lobes: specular 1, clearcoat 1/4 -> probabilities 4/5 and 1/5
        ax, ay = _aniso(0.5, 0.0)
        want_pdf = 0.8 * G1(wo, ax, ay) * D_tr(h, ax, ay) / (4 * wo[2]) + 0.2 * D_tr(h, a, a) / (4 * (wo @ h))
        assert pdf1 == pytest.approx(want_pdf, rel=3e-4)


def _hemisphere_integral(m, wo, n_theta=160, n_phi=320):
    """Midpoint quadrature of integral f(wo, wi) cos dwi over the upper hemisphere."""
    th = (np.arange(n_theta) + 0.5) * (PI / 2 / n_theta)
    ph = (np.arange(n_phi) + 0.5) * (2 * PI / n_phi)
    tot = np.zeros(3)
    for t in th:
        st, ct = math.sin(t), math.cos(t)
        for p in ph:
            wi = np.array([st * math.cos(p), st * math.sin(p), ct])
            f, _, _ = oracle.eval_bsdf(m, N, wo, wi)
            tot += f * ct * st
    return tot * (PI / 2 / n_theta) * (2 * PI / n_phi)


@pytest.mark.parametrize("cfg", [
    dict(roughness=0.5, metallic=0.0),
    dict(roughness=0.35, metallic=1.0, anisotropy=0.6),
    dict(roughness=0.6, metallic=0.4, specular_tint=0.5),
], ids=["dielectric", "metal_aniso", "mixed"])
def test_principled_sampler_density_is_bsdf_pdf(cfg):
    """If sample_bsdf draws wi with density bsdf_pdf(wi), then
    E[f(wi) cos / bsdf_pdf(wi)] over its samples equals the hemisphere
    integral of f cos.  A lobe probability or a sampling pdf that disagrees
    with bsdf_pdf shifts the mean.  (Clearcoat is left out: the reference's
    clearcoat sampler builds its half vector from a zero cos_theta argument,
    principled.glsl:143, so it does not sample its own pdf.)"""
    m = _mat(bsdf_type=32, bsdf_props=8 | 4, ior=IOR, albedo=(0.8, 0.6, 0.4), **cfg)
    wo = _unit([0.4, -0.2, 0.8])
    ref = _hemisphere_integral(m, wo)
    rng = np.random.default_rng(11)
    n = 40000
    acc = np.zeros(3)
    acc2 = np.zeros(3)
    for r in rng.random((n, 3)):
        wi, f, pdf, c = oracle.sample_bsdf(m, N, wo, r)
        if pdf <= 0 or wi[2] <= 0:
            continue
        fe, _, _ = oracle.eval_bsdf(m, N, wo, wi)
        p = oracle.bsdf_pdf(m, N, wo, wi)
        v = fe * wi[2] / p
        acc += v
        acc2 += v * v
    mean = acc / n
    se = np.sqrt(np.maximum(acc2 / n - mean ** 2, 0) / n)
    assert np.all(np.abs(mean - ref) <= 4 * se + 2e-3 * ref), (mean, ref, se)


def test_principled_single_lobe_sampled_pdf_is_bsdf_pdf():
    """metallic = 1 has one lobe: the pdf returned with a sample is bsdf_pdf."""
    m = _mat(bsdf_type=32, bsdf_props=8 | 4, roughness=0.4, metallic=1.0, ior=IOR, albedo=(0.9, 0.7, 0.5))
    wo = _unit([0.4, -0.2, 0.8])
    rng = np.random.default_rng(2)
    for r in rng.random((500, 3)):
        wi, f, pdf, c = oracle.sample_bsdf(m, N, wo, r)
        if pdf <= 0 or not np.any(f):  # a reflection below the horizon returns f = 0
            continue
        assert oracle.bsdf_pdf(m, N, wo, wi) == pytest.approx(pdf, rel=2e-3)
        fe, _, _ = oracle.eval_bsdf(m, N, wo, wi)
        assert fe == pytest.approx(f, rel=1e-4)


# --------------------------------------------------------- point lights ----
def _point_plane(lights, rho=0.6):
    from scenes import DescScene, quad
    s = DescScene()
    m = s.material(bsdf_type=1, bsdf_props=9, albedo=(rho,) * 3)
    s.mesh(quad([-50, 0, -50], [-50, 0, 50], [50, 0, 50], [50, 0, -50]), m)
    for pos, I in lights:
        s.analytic.append(dict(pos=pos, to=(0, 0, 0), L=(I, I, I), flags=4 | 16 | 32))
    return s


def test_point_light_Li_closed_form():
    """An isotropic point light of intensity I seen from p: direction to the
    light, distance d, Le = I and pdf_w = d^2, so the NEE estimate of the
    irradiance Le cos / pdf_w is I cos / d^2."""
    s = _point_plane([((1.0, 3.0, -2.0), 7.0)])
    d = s.build()
    p = np.array([0.5, 0.0, 0.5])
    o = oracle.sample_light_Li(d, [0.3, 0.1, 0.2, 0.9], p)
    Le, wi, wl, pw, pa = o[:3], o[3:6], o[6], o[7], o[8]
    v = np.array([1.0, 3.0, -2.0]) - p
    dist = np.linalg.norm(v)
    assert np.allclose(wi, v / dist, atol=1e-6)
    assert wl == pytest.approx(dist, rel=1e-6)
    assert np.allclose(Le, 7.0)
    assert pw == pytest.approx(dist ** 2, rel=1e-5) and pa == 1.0
    assert Le[0] * wi[1] / pw == pytest.approx(7.0 * (3.0 / dist) / dist ** 2, rel=1e-5)


def test_point_light_Le_is_isotropic():
    """Emission sampling of a point light: origin at the light, directions
    uniform on the sphere (mean 0, <z^2> = 1/3), pdf_dir = 1/(4 pi), pdf_pos
    = 1 per light."""
    s = _point_plane([((0.0, 2.0, 0.0), 5.0)])
    d = s.build()
    rng = np.random.default_rng(4)
    dirs = []
    for r in rng.random((4000, 6)):
        o = oracle.sample_light_Le(d, r)
        Le, pos, wi, pdf_pos, pdf_dir = o[:3], o[3:6], o[6:9], o[13], o[14]
        assert np.allclose(pos, [0, 2, 0]) and np.allclose(Le, 5.0)
        assert pdf_dir == pytest.approx(1 / (4 * PI), rel=1e-6) and pdf_pos == pytest.approx(1.0)
        dirs.append(wi)
    dirs = np.array(dirs)
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1, atol=1e-5)
    assert np.all(np.abs(dirs.mean(0)) < 0.05)
    assert np.allclose((dirs ** 2).mean(0), 1 / 3, atol=0.03)


def test_point_lit_plane_path_closed_form():
    """Two point lights over a Lambert plane, Path at depth 2 (direct light
    only): the NEE picks one light with probability 1/2 and weights by
    light_triangle_count = 2 (path.rgen:76-78; each analytic light counts 1),
    so the mean over frames is rho/pi sum_k I_k cos_k / d_k^2."""
    rho = 0.6
    lights = [((0.0, 3.0, 0.0), 20.0), ((2.0, 1.5, -1.0), 6.0)]
    s = _point_plane(lights, rho)
    acc = oracle.Accel(s.build())
    from scenes import camera
    cam = camera([0.0, 10.0, 0.0], [0.0, -1.0, 0.0], 1.0)
    img, _, _, _ = acc.render_path(cam, 9, 9, 2, frames=256)
    want = 0.0
    for pos, I in lights:
        v = np.array(pos)
        dist = np.linalg.norm(v)
        want += rho / PI * I * (v[1] / dist) / dist ** 2
    assert img[4, 4, :3] == pytest.approx([want] * 3, rel=0.03)


@pytest.mark.parametrize("integrator", ["bdpt", "vcm"])
def test_point_lit_plane_bdpt_vcm_closed_form(integrator):
    """One point light of intensity I at height h over a Lambert plane: the
    radiance of the point below it is rho/pi I / h^2 under BDPT (NEE and t = 1
    splats, MIS-weighted) and under the corrected VCM estimator."""
    rho, I, h = 0.7, 20.0, 3.0
    s = _point_plane([((0.0, h, 0.0), I)], rho)
    acc = oracle.Accel(s.build())
    from scenes import camera
    cam = camera([0.0, 10.0, 0.0], [0.0, -1.0, 0.0], 1.0)
    if integrator == "bdpt":
        img, _, _, _ = acc.render_bdpt(cam, 9, 9, 2, frames=16)
    else:
        img, _, _, _ = acc.render_vcm(cam, 9, 9, 2, frames=16, corrected=True, use_vm=False)
    want = rho / PI * I / h ** 2
    assert img[4, 4, 0] == pytest.approx(want, rel=0.02)
