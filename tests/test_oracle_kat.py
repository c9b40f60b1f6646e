"""Pins the CPU oracle: known answers, closed forms and invariants of the
restated reference functions (RNG, ray offsets, sampling, Fresnel, GGX,
BSDFs, light sampling).  CPU only."""
import ctypes as C
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from lumen_amd import _abi

GOLD = Path(__file__).parent / "golden"


# ---------------------------------------------------------------- RNG ----
def test_pcg4d_matches_published_listing_fixture():
    g = json.loads((GOLD / "pcg4d.json").read_text())
    out = oracle.pcg4d(np.array(g["inputs"], np.uint32))
    assert np.array_equal(out, np.array(g["outputs"], np.uint32))


def test_rand_sequence_fixture_and_counter_semantics():
    g = json.loads((GOLD / "pcg4d.json").read_text())
    for s in g["rand_sequences"]:
        got = oracle.rand_seq(s["x"], s["y"], s["frame"], len(s["rand"]))
        assert np.array_equal(got, np.array(s["rand"], np.float32))


def test_uint_to_float_closed_form():
    L = oracle.lib()
    assert L.ora_uint_to_float(0) == 0.0
    assert L.ora_uint_to_float(0xFFFFFFFF) == np.float32(1.0 - 2.0 ** -23)
    assert L.ora_uint_to_float(1 << 31) == 0.5
    assert L.ora_uint_to_float(511) == 0.0  # low 9 bits are dropped


def test_rand_uniformity_chi2():
    x = np.concatenate([oracle.rand_seq(i, 7, 3, 64) for i in range(256)])
    assert x.min() >= 0.0 and x.max() < 1.0
    hist, _ = np.histogram(x, bins=32, range=(0, 1))
    e = len(x) / 32
    chi2 = ((hist - e) ** 2 / e).sum()
    assert chi2 < 70  # 31 dof, p ~ 1e-4


# -------------------------------------------------------------- offsets ----
def test_offset_ray_integer_ulp_push():
    # |p| >= 1/32: the float's bit pattern moves by int(256 * n) ulps (RT Gems ch. 6)
    p = np.array([1.0, -2.0, 0.5], np.float32)
    n = np.array([1.0, 0.5, -0.25], np.float32)
    got = oracle.offset_ray(p, n)
    bits = p.view(np.int32).copy()
    bits += np.array([256, -128, -64], np.int32)  # sign of p flips the step
    assert np.array_equal(got.view(np.int32), bits)


def test_offset_ray_near_origin_uses_float_scale():
    p = np.array([0.01, -0.02, 0.0], np.float32)
    n = np.array([0.0, 1.0, -1.0], np.float32)
    got = oracle.offset_ray(p, n)
    want = p + np.float32(1.0 / 65536.0) * n
    assert np.array_equal(got, want)


def test_offset_ray2_closed_form():
    p = np.array([3.0, -4.0, 5.0], np.float32)
    n = np.array([0.0, 0.0, 1.0], np.float32)
    assert np.array_equal(oracle.offset_ray(p, n, which=2), p + np.float32(2.0 / 65536.0) * n)


# ------------------------------------------------------------- sampling ----
def test_concentric_disk_known_points_and_area_preservation():
    assert np.allclose(oracle.concentric_disk([1.0, 0.5]), [1.0, 0.0], atol=1e-7)
    assert np.allclose(oracle.concentric_disk([0.5, 1.0]), [0.0, 1.0], atol=1e-6)
    assert np.allclose(oracle.concentric_disk([0.5, 0.5]), [0.0, 0.0])
    rng = np.random.default_rng(1)
    pts = np.array([oracle.concentric_disk(r) for r in rng.random((4000, 2))])
    r2 = (pts ** 2).sum(1)
    assert r2.max() <= 1.0 + 1e-6
    # uniform on the disk <=> r^2 uniform on [0, 1]
    hist, _ = np.histogram(r2, bins=10, range=(0, 1))
    assert np.abs(hist / 400.0 - 1.0).max() < 0.2


@pytest.mark.parametrize("n", [[0, 0, 1], [0, 0, -1], [1, 0, 0], [0.3, -0.4, 0.866], [-0.6, 0.0, -0.8]])
def test_branchless_onb_is_orthonormal_right_handed(n):
    n = np.array(n, np.float32)
    n /= np.linalg.norm(n)
    b1, b2 = oracle.onb(n)
    M = np.stack([b1, b2, n])
    assert np.allclose(M @ M.T, np.eye(3), atol=1e-6)
    assert np.allclose(np.cross(b1, b2), n, atol=1e-6)


# --------------------------------------------------------------- Fresnel ----
@pytest.mark.parametrize("eta", [1.33, 1.5, 2.4])
def test_fresnel_dielectric_closed_forms(eta):
    L = oracle.lib()
    r0 = ((eta - 1) / (eta + 1)) ** 2
    assert L.ora_fresnel_dielectric(1.0, eta, 1) == pytest.approx(r0, rel=1e-5)
    assert L.ora_fresnel_dielectric(1e-7, eta, 1) == pytest.approx(1.0, abs=1e-3)
    # from inside: total internal reflection beyond the critical angle
    crit = math.sqrt(1 - 1 / eta ** 2)
    assert L.ora_fresnel_dielectric(crit * 0.9, eta, 0) == 1.0
    assert L.ora_fresnel_dielectric(1.0, eta, 0) == pytest.approx(r0, rel=1e-5)


def test_fresnel_conductor_reduces_to_dielectric_without_absorption():
    L = oracle.lib()
    for c in np.linspace(0.05, 1.0, 12):
        assert L.ora_fresnel_conductor(float(c), 1.5, 0.0) == pytest.approx(
            L.ora_fresnel_dielectric(float(c), 1.5, 1), rel=2e-4, abs=1e-6)


@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.6, 1.0])
def test_ggx_isotropic_normalisation(alpha):
    # integral over the hemisphere of D(h) cos(theta_h) = 1
    L = oracle.lib()
    th = np.linspace(0, np.pi / 2, 20001)[:-1] + np.pi / 4 / 20000
    vals = np.array([L.ora_D_ggx_iso(alpha * alpha, float(np.cos(t))) for t in th])
    integ = (vals * np.cos(th) * np.sin(th)).sum() * (np.pi / 2 / 20000) * 2 * np.pi
    assert integ == pytest.approx(1.0, rel=2e-3)


# ----------------------------------------------------------------- BSDFs ----
def _mat(**kw):
    m = _abi.lt_material()
    m.texture_id = -1
    for k, v in kw.items():
        if isinstance(v, (tuple, list)):
            getattr(m, k)[:] = list(v)
        else:
            setattr(m, k, v)
    return m


DIFF = dict(bsdf_type=1, bsdf_props=1 | 8, albedo=(0.5, 0.6, 0.7))
ROUGH_COND = dict(bsdf_type=16, bsdf_props=8 | 4, roughness=0.4, albedo=(0.2, 0.9, 1.1), k=(3.9, 2.4, 2.2))
PRINC = dict(bsdf_type=32, bsdf_props=8 | 4, roughness=0.5, metallic=0.0, ior=1.5, albedo=(0.8, 0.5, 0.3),
             sheen_tint=0.5, clearcoat_gloss=1.0)
PRINC_METAL = dict(PRINC, metallic=1.0, roughness=0.3)
ROUGH_DIEL_R = dict(bsdf_type=8, bsdf_props=8 | 4, roughness=0.5, ior=1.5, albedo=(1, 1, 1))


def test_lambert_eval_and_pdf_closed_form():
    m = _mat(**DIFF)
    n = np.array([0, 0, 1.0], np.float32)
    wo = np.array([0.3, 0.1, 0.9], np.float32)
    wo /= np.linalg.norm(wo)
    wi = np.array([-0.5, 0.2, 0.7], np.float32)
    wi /= np.linalg.norm(wi)
    f, pdf, pdf_rev = oracle.eval_bsdf(m, n, wo, wi)
    assert np.allclose(f, np.array(DIFF["albedo"]) / np.pi, rtol=1e-5)
    assert pdf == pytest.approx(wi[2] / np.pi, rel=1e-5)
    f2, _, _ = oracle.eval_bsdf(m, n, wo, -wi)
    assert np.all(f2 == 0)


def _sample_stats(m, wo, n_s, count=4000, seed=0):
    rng = np.random.default_rng(seed)
    wsum = np.zeros(3)
    bad = 0
    for r in rng.random((count, 3)):
        wi, f, pdf, c = oracle.sample_bsdf(m, n_s, wo, r)
        if pdf <= 0 or not np.any(f):
            continue
        wsum += f * abs(c) / pdf
        p2 = oracle.bsdf_pdf(m, n_s, wo, wi)
        if not (abs(p2 - pdf) <= 2e-3 * max(pdf, 1e-3)):
            bad += 1
    return wsum / count, bad


@pytest.mark.parametrize("cfg", [DIFF, ROUGH_COND, PRINC, PRINC_METAL], ids=["lambert", "conductor", "principled",
                                                                             "principled_metal"])
def test_sampled_pdf_equals_pdf_and_furnace_bound(cfg):
    m = _mat(**cfg)
    n = np.array([0, 0, 1.0], np.float32)
    wo = np.array([0.4, -0.2, 0.8], np.float32)
    wo /= np.linalg.norm(wo)
    est, bad = _sample_stats(m, wo, n)
    # white-furnace style bound on the albedo estimate of a passive BSDF
    assert np.all(est <= 1.0 + 0.05), est
    # for every sampled direction of a single-lobe BSDF the returned pdf is bsdf_pdf's
    if cfg is not PRINC and cfg is not PRINC_METAL:
        assert bad == 0


@pytest.mark.parametrize("cfg", [DIFF, ROUGH_COND, ROUGH_DIEL_R, PRINC], ids=["lambert", "conductor", "dielectric",
                                                                              "principled"])
def test_pdf_integrates_to_at_most_one(cfg):
    m = _mat(**cfg)
    n = np.array([0, 0, 1.0], np.float32)
    wo = np.array([0.2, 0.1, 0.97], np.float32)
    wo /= np.linalg.norm(wo)
    rng = np.random.default_rng(3)
    d = rng.normal(size=(20000, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    vals = np.array([oracle.bsdf_pdf(m, n, wo, w) for w in d])
    integ = vals.mean() * 4 * np.pi
    assert integ <= 1.05 and integ > 0.3, integ


def test_lambert_furnace_is_albedo():
    m = _mat(**DIFF)
    est, _ = _sample_stats(m, np.array([0, 0, 1.0], np.float32), np.array([0, 0, 1.0], np.float32), 3000)
    assert np.allclose(est, DIFF["albedo"], rtol=1e-4)


def test_smooth_dielectric_energy_split_and_directions():
    m = _mat(bsdf_type=8, bsdf_props=8 | 16 | 2, ior=1.5, albedo=(1, 1, 1))
    n = np.array([0, 0, 1.0], np.float32)
    wo = np.array([0.0, 0.6, 0.8], np.float32)
    F = oracle.lib().ora_fresnel_dielectric(0.8, 1.5, 1)
    # reflection branch (rand < F): mirror direction, weight f*cos/pdf = 1
    wi, f, pdf, c = oracle.sample_bsdf(m, n, wo, [F * 0.5, 0.3, 0.3])
    assert np.allclose(wi, [0.0, -0.6, 0.8], atol=1e-6)
    assert np.allclose(f * abs(c) / pdf, 1.0, rtol=1e-5)
    # refraction branch: Snell's law, radiance scaled by 1/eta^2 (mode 1)
    wi, f, pdf, c = oracle.sample_bsdf(m, n, wo, [0.99, 0.3, 0.3])
    sin_t = np.hypot(wi[0], wi[1])
    assert sin_t == pytest.approx(0.6 / 1.5, rel=1e-5) and wi[2] < 0
    assert np.allclose(f * abs(c) / pdf, 1.0 / 1.5 ** 2, rtol=1e-5)


# ---------------------------------------------------------------- lights ----
def test_triangle_light_pdf_is_inverse_area_and_point_on_light():
    from scenes import DescScene, quad
    s = DescScene()
    lm = s.material(bsdf_type=1, bsdf_props=9, albedo=(1, 1, 1), emissive_factor=(5, 5, 5))
    s.mesh(quad([0, 2, 0], [1, 2, 0], [1, 2, 0.5], [0, 2, 0.5]), lm)  # two triangles of area 0.25
    d = s.build()
    rng = np.random.default_rng(0)
    for r in rng.random((200, 4)):
        o = oracle.sample_light_Li(d, r, [0.3, 0.0, 0.2])
        Le, wi, wl, pw, pa = o[:3], o[3:6], o[6], o[7], o[8]
        assert pa == pytest.approx(4.0, rel=1e-5)  # 1 / area of the sampled triangle
        hit = np.array([0.3, 0.0, 0.2]) + wi * wl
        assert hit[1] == pytest.approx(2.0, abs=1e-5) and 0 <= hit[0] <= 1 and 0 <= hit[2] <= 0.5
        assert pw == pytest.approx(pa * wl * wl / abs(wi[1]), rel=1e-4)
        assert np.allclose(Le, 5.0)


@pytest.mark.parametrize("angle,expect", [(0.0, 1.0), (20.0, 1.0), (35.0, 0.0)])
def test_spot_falloff_closed_form(angle, expect):
    from scenes import DescScene, quad
    s = DescScene()
    w = s.material(bsdf_type=1, bsdf_props=9, albedo=(1, 1, 1))
    s.mesh(quad([-5, 0, -5], [5, 0, -5], [5, 0, 5], [-5, 0, 5]), w)
    s.analytic.append(dict(pos=(0, 4, 0), to=(0, 0, 0), L=(10, 10, 10), flags=1 | 16 | 32))
    d = s.build()
    x = 4 * math.tan(math.radians(angle))
    o = oracle.sample_light_Li(d, [0.0, 0.5, 0.5, 0.5], [x, 0, 0])
    assert o[0] == pytest.approx(10 * expect, abs=1e-5)
    # between 25 and 30 degrees: ((cos - cos30) / (cos25 - cos30))^4
    x = 4 * math.tan(math.radians(27.5))
    o = oracle.sample_light_Li(d, [0.0, 0.5, 0.5, 0.5], [x, 0, 0])
    c = math.cos(math.radians(27.5))
    want = ((c - math.cos(math.pi / 6)) / (math.cos(math.radians(25)) - math.cos(math.pi / 6))) ** 4
    assert o[0] == pytest.approx(10 * want, rel=1e-3)
