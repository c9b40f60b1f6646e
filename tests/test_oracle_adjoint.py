"""The differentiable Path estimator and its path-replay adjoint in the oracle
(oracle/oracle_adjoint.cpp).  No reference counterpart exists; the adjoint
is pinned by
  - the BSDF derivatives against central differences of eval_bsdf (itself
    pinned by the closed-form BSDF tests);
  - the forward estimator against the Path integrator's own frames (one
    sample = one Path frame while max_depth <= 5 keeps roulette out);
  - texel gradients against central differences (h = 1e-3) of the forward
    estimator evaluated in float64, on 64 random texels, within 1e-3
    relative: albedo directly (paths do not depend on albedo, so the
    estimator is a polynomial in it), roughness with sampling held at the
    base texels (the detached-sampling estimator the adjoint
    differentiates).
"""
import ctypes as C

import numpy as np
import pytest

import lumen_amd as la
import oracle
from lumen_amd import _abi


def _textures(sc):
    d = sc.desc
    out = []
    for t in range(d.num_textures):
        tex = d.textures[t]
        a = np.ctypeslib.as_array(tex.data, shape=(tex.height, tex.width, 4)).copy()
        out.append(a)
    return out


def _setup(W=24, H=24, wall_tris=128, tex_res=16, seed=5):
    sc = la.Scene.procedural("diff_wall", seed=seed, wall_tris=wall_tris, tex_res=tex_res)
    tex = _textures(sc)
    alb = tex[0][..., :3].astype(np.float32)
    rough = tex[1][..., 0].astype(np.float32)
    return sc, sc.camera(W, H), oracle.Accel(sc.desc), alb, rough


def _principled(albedo, roughness):
    m = _abi.lt_material()
    m.albedo[:] = list(albedo)
    m.bsdf_type = 32
    m.bsdf_props = 2 | (8 if roughness > 0.08 else 4)
    m.ior = 1.5
    m.roughness = roughness
    m.sheen_tint = 0.5
    m.clearcoat_gloss = 1.0
    m.texture_id = -1
    return m


def _hemi(rng, n):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v[:, 2] = np.abs(v[:, 2]) + 0.02
    return v / np.linalg.norm(v, axis=1, keepdims=True)


@pytest.mark.parametrize("kind", ["principled", "lambert"])
def test_bsdf_derivatives_match_central_differences(kind):
    rng = np.random.default_rng(3)
    n_s = np.array([0.0, 0.0, 1.0], np.float32)
    wos, wis = _hemi(rng, 300), _hemi(rng, 300)
    h = 1e-3
    checked = 0
    for i in range(300):
        albedo = rng.uniform(0.1, 0.9, 3)
        r = float(rng.uniform(0.12, 0.95))
        if kind == "principled":
            m = _principled(albedo, r)
        else:
            m = _abi.lt_material()
            m.albedo[:] = list(albedo)
            m.bsdf_type, m.bsdf_props, m.texture_id = 1, 1 | 8, -1
        f, dfa, dfr = oracle.bsdf_eval_d(m, n_s, wos[i], wis[i])
        for c in range(3):
            mp, mm = _abi.lt_material.from_buffer_copy(m), _abi.lt_material.from_buffer_copy(m)
            mp.albedo[c] += h
            mm.albedo[c] -= h
            fd = (oracle.eval_bsdf(mp, n_s, wos[i], wis[i])[0][c].astype(np.float64)
                  - oracle.eval_bsdf(mm, n_s, wos[i], wis[i])[0][c]) / (2 * h)
            assert dfa[c] == pytest.approx(fd, rel=2e-3, abs=2e-4), (i, c)
        mp, mm = _abi.lt_material.from_buffer_copy(m), _abi.lt_material.from_buffer_copy(m)
        mp.roughness += h
        mm.roughness -= h
        fd = (oracle.eval_bsdf(mp, n_s, wos[i], wis[i])[0].astype(np.float64)
              - oracle.eval_bsdf(mm, n_s, wos[i], wis[i])[0]) / (2 * h)
        scale = max(1.0, float(np.abs(f).max()))
        assert np.allclose(dfr, fd, rtol=1e-2, atol=2e-3 * scale), (i, dfr, fd)
        checked += 1
    assert checked == 300
    if kind == "lambert":
        assert np.allclose(dfa, 1 / np.float32(3.14159265359)) and np.all(dfr == 0)


def test_delta_lobe_has_no_roughness_derivative_from_reflection():
    """alpha = r^2 <= 0.0064 is the reference's mirror lobe (eval returns
    only the diffuse term): the roughness derivative is the diffuse one."""
    m = _principled([0.5, 0.5, 0.5], 0.05)
    n_s = np.array([0.0, 0.0, 1.0], np.float32)
    wo = np.array([0.3, 0.1, 0.9], np.float32)
    wo /= np.linalg.norm(wo)
    wi = np.array([-0.3, -0.1, 0.9], np.float32)
    wi /= np.linalg.norm(wi)
    f, dfa, dfr = oracle.bsdf_eval_d(m, n_s, wo, wi)
    mp, mm = _abi.lt_material.from_buffer_copy(m), _abi.lt_material.from_buffer_copy(m)
    mp.roughness += 1e-3
    mm.roughness -= 1e-3
    fd = (oracle.eval_bsdf(mp, n_s, wo, wi)[0].astype(np.float64) - oracle.eval_bsdf(mm, n_s, wo, wi)[0]) / 2e-3
    assert np.allclose(dfr, fd, rtol=1e-2, atol=1e-4)


def test_forward_is_the_path_integrator_without_roulette():
    sc, cam, acc, alb, rough = _setup()
    W = H = 24
    for seed in (0, 9):
        img = acc.adjoint(cam, W, H, 5, 1, albedo=alb[None], rough=rough[None], albedo_id=0, rough_id=1, seed=seed)
        ref, _, _, _ = acc.render_path(cam, W, H, 5, seed=seed)
        assert np.array_equal(img[0], ref[..., :3]), seed
    img = acc.adjoint(cam, W, H, 5, 4, albedo=alb[None], rough=rough[None], albedo_id=0, rough_id=1, seed=9)
    ref, _, _, _ = acc.render_path(cam, W, H, 5, frames=4, seed=9)
    assert np.allclose(img[0], ref[..., :3], rtol=1e-5, atol=1e-6)


def _loss_weights(rng, V, H, W):
    return rng.normal(size=(V, H, W, 3)).astype(np.float32)


def _fd_texels(rng, grad, n):
    """n random entries among those the image depends on."""
    flat = np.flatnonzero(np.abs(grad) > 1e-6 * np.abs(grad).max())
    return rng.choice(flat, size=min(n, len(flat)), replace=False)


# The adjoint's texel gradients against central differences with h = 1e-3 of
# the same estimator evaluated in float64 (oracle.Accel64): in float32 the
# difference of two images is swamped by rounding (about 1e-7 of a pixel, or
# 5e-5 of it after dividing by 2h), which reaches tens of percent of the
# gradients of weakly seen texels.  In float64 the central difference is the
# derivative up to O(h^2): the bar is 1e-3 relative on every texel, plus a
# floor of 1e-6 of the largest gradient for float32 summation in the adjoint.
FD_REL = 1e-3
FD_FLOOR = 1e-6


def test_albedo_gradient_matches_float64_central_differences():
    sc, cam, acc, alb, rough = _setup()
    a64 = oracle.Accel64(sc.desc)
    W = H = 24
    D, spp, h = 4, 2, 1e-3
    rng = np.random.default_rng(11)
    albs = np.stack([alb, np.clip(alb + rng.uniform(-0.1, 0.1, alb.shape), 0.05, 0.95)]).astype(np.float32)
    roughs = np.stack([rough, rough]).astype(np.float32)
    w = _loss_weights(rng, 2, H, W)
    ga, gr = acc.adjoint(cam, W, H, D, spp, albedo=albs, rough=roughs, albedo_id=0, rough_id=1, seed=2, d_images=w)
    assert ga.shape == albs.shape and gr.shape == roughs.shape
    picks = _fd_texels(rng, ga, 64)
    assert len(picks) == 64
    errs = []
    for flat in picks:
        v, ty, tx, c = np.unravel_index(flat, ga.shape)
        ap, am = albs.astype(np.float64), albs.astype(np.float64)
        ap[v, ty, tx, c] += h
        am[v, ty, tx, c] -= h
        kw = dict(rough=roughs[v:v + 1], albedo_id=0, rough_id=1, seed=2)
        ip = a64.adjoint(cam, W, H, D, spp, albedo=ap[v:v + 1], **kw)
        im = a64.adjoint(cam, W, H, D, spp, albedo=am[v:v + 1], **kw)
        fd = float(np.sum(w[v] * (ip[0] - im[0])) / (2 * h))
        errs.append(abs(ga.flat[flat] - fd) / abs(fd))
        assert abs(ga.flat[flat] - fd) <= FD_REL * abs(fd) + FD_FLOOR * np.abs(ga).max(), (v, ty, tx, c, fd)
    print(f"albedo: max rel error {max(errs):.2e} over 64 texels")


def test_roughness_gradient_matches_detached_float64_central_differences():
    """Roughness moves the sampled directions; the adjoint differentiates the
    detached-sampling estimator (sampling at the base texels, BSDF values at
    the perturbed ones), whose central differences it must match."""
    sc, cam, acc, alb, rough = _setup()
    a64 = oracle.Accel64(sc.desc)
    W = H = 24
    D, spp, h = 4, 2, 1e-3
    rng = np.random.default_rng(12)
    w = _loss_weights(rng, 1, H, W)
    ga, gr = acc.adjoint(cam, W, H, D, spp, albedo=alb[None], rough=rough[None], albedo_id=0, rough_id=1, seed=4,
                         d_images=w)
    picks = _fd_texels(rng, gr, 64)
    assert len(picks) == 64
    errs = []
    for flat in picks:
        _, ty, tx = np.unravel_index(flat, gr.shape)
        rp, rm = rough.astype(np.float64), rough.astype(np.float64)
        rp[ty, tx] += h
        rm[ty, tx] -= h
        kw = dict(albedo=alb[None], rough=rough[None], albedo_id=0, rough_id=1, seed=4)
        ip = a64.adjoint(cam, W, H, D, spp, eval_rough=rp[None], **kw)
        im = a64.adjoint(cam, W, H, D, spp, eval_rough=rm[None], **kw)
        fd = float(np.sum(w[0] * (ip[0] - im[0])) / (2 * h))
        errs.append(abs(gr.flat[flat] - fd) / abs(fd))
        assert abs(gr.flat[flat] - fd) <= FD_REL * abs(fd) + FD_FLOOR * np.abs(gr).max(), (ty, tx, fd)
    print(f"roughness: max rel error {max(errs):.2e} over 64 texels")


def test_float64_forward_follows_the_float32_paths():
    """The float64 estimator takes the float32 estimator's paths: images agree
    to float32 rounding (a flipped decision would differ by O(1))."""
    sc, cam, acc, alb, rough = _setup()
    a64 = oracle.Accel64(sc.desc)
    kw = dict(albedo=alb[None], rough=rough[None], albedo_id=0, rough_id=1, seed=6)
    i32 = acc.adjoint(cam, 24, 24, 5, 2, **kw)
    i64 = a64.adjoint(cam, 24, 24, 5, 2, **kw)
    assert np.allclose(i32, i64, rtol=1e-4, atol=1e-6 * np.abs(i64).max())


def test_frozen_sampling_at_base_texels_is_the_forward():
    sc, cam, acc, alb, rough = _setup()
    kw = dict(albedo=alb[None], rough=rough[None], albedo_id=0, rough_id=1, seed=4)
    a = acc.adjoint(cam, 24, 24, 4, 2, **kw)
    b = acc.adjoint(cam, 24, 24, 4, 2, eval_rough=rough[None], eval_albedo=alb[None], **kw)
    assert np.allclose(a, b, rtol=1e-5, atol=1e-6)


def test_variants_are_independent():
    sc, cam, acc, alb, rough = _setup(W=16, H=16)
    rng = np.random.default_rng(1)
    albs = np.stack([alb, alb * 0.8]).astype(np.float32)
    roughs = np.stack([rough, np.clip(rough + 0.1, 0, 1)]).astype(np.float32)
    w = _loss_weights(rng, 2, 16, 16)
    kw = dict(albedo_id=0, rough_id=1, seed=3)
    img = acc.adjoint(cam, 16, 16, 4, 2, albedo=albs, rough=roughs, **kw)
    ga, gr = acc.adjoint(cam, 16, 16, 4, 2, albedo=albs, rough=roughs, d_images=w, **kw)
    for v in range(2):
        i1 = acc.adjoint(cam, 16, 16, 4, 2, albedo=albs[v:v + 1], rough=roughs[v:v + 1], **kw)
        g1a, g1r = acc.adjoint(cam, 16, 16, 4, 2, albedo=albs[v:v + 1], rough=roughs[v:v + 1], d_images=w[v:v + 1],
                               **kw)
        assert np.array_equal(img[v], i1[0])
        assert np.allclose(ga[v], g1a[0], rtol=1e-12, atol=0) and np.allclose(gr[v], g1r[0], rtol=1e-12, atol=0)
    assert not np.allclose(img[0], img[1])
