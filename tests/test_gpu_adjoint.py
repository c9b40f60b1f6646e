"""GPU path-replay adjoint (lt_adjoint_*, csrc/adjoint.hip) against the
oracle's differentiable Path estimator and its adjoint
(oracle/oracle_adjoint.cpp), plus central differences on the GPU itself.

Same scene, camera, seed and texels: each sample replays the oracle's pcg4d
stream, so the images agree to float tolerance (the Path replay bar:
|gpu - cpu| <= 1e-3 (1 + |cpu|) on >= 99.5% of the pixels, relative RMSE
<= 1%).  Gradients are sums of many float atomics on the GPU and of doubles
on the host; they must agree to a relative L2 error of 1e-4 with >= 99.9% of
the texels within 1e-4 * max|g| + 1e-3 |g| (measured: about 1e-6).
"""
import numpy as np
import pytest

import lumen_amd as la
import oracle
from test_gpu_bdpt import _compare

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _scene(tex_res=16, wall_tris=512, seed=5):
    sc = la.Scene.procedural("diff_wall", seed=seed, wall_tris=wall_tris, tex_res=tex_res)
    d = sc.desc
    tex = []
    for t in range(d.num_textures):
        tx = d.textures[t]
        tex.append(np.ctypeslib.as_array(tx.data, shape=(tx.height, tx.width, 4)).copy())
    return sc, tex[0][..., :3].astype(np.float32), tex[1][..., 0].astype(np.float32)


def _variants(alb, rough, V, rng):
    albs = [alb] + [np.clip(alb + rng.uniform(-0.15, 0.15, alb.shape), 0.05, 0.95) for _ in range(V - 1)]
    roughs = [rough] + [np.clip(rough + rng.uniform(-0.2, 0.2, rough.shape), 0.1, 0.95) for _ in range(V - 1)]
    return np.stack(albs).astype(np.float32), np.stack(roughs).astype(np.float32)


def _cmp_grad(g, o, tag):
    g = np.asarray(g, np.float64)
    o = np.asarray(o, np.float64)
    assert np.all(np.isfinite(g))
    err = np.abs(g - o)
    ok = err <= 1e-4 * np.abs(o).max() + 1e-3 * np.abs(o)
    rel = np.linalg.norm(g - o) / max(np.linalg.norm(o), 1e-30)
    print(f"[gpu] {tag}: texels matching {ok.mean():.4f}, rel L2 {rel:.2e}", flush=True)
    assert ok.mean() >= 0.999
    assert rel <= 1e-4


@pytest.mark.parametrize("W,H,V,spp,depth", [(24, 24, 2, 2, 4), (40, 32, 3, 2, 6)])
def test_adjoint_replays_oracle(gpu_ctx, W, H, V, spp, depth):
    print(f"[gpu] adjoint replay {W}x{H} V={V} spp={spp} depth={depth}", flush=True)
    from lumen_amd.adjoint import Adjoint
    sc, alb, rough = _scene()
    rng = np.random.default_rng(W + V)
    albs, roughs = _variants(alb, rough, V, rng)
    w = rng.normal(size=(V, H, W, 3)).astype(np.float32)
    adj = Adjoint(sc, gpu_ctx, W, H, spp=spp, variants=V, albedo_texture=0, roughness_texture=1, max_depth=depth)
    ta, tr = torch.from_numpy(albs).cuda(), torch.from_numpy(roughs).cuda()
    img, cache = adj.forward(ta, tr, seed=7)
    ga, gr = adj.backward(ta, tr, torch.from_numpy(w).cuda(), seed=7, cache=cache)
    ga2, gr2 = adj.backward(ta, tr, torch.from_numpy(w).cuda(), seed=7, cache=None)
    torch.cuda.synchronize()
    acc = oracle.Accel(sc.desc)
    o_img = acc.adjoint(adj.camera, W, H, depth, spp, albedo=albs, rough=roughs, albedo_id=0, rough_id=1, seed=7)
    o_ga, o_gr = acc.adjoint(adj.camera, W, H, depth, spp, albedo=albs, rough=roughs, albedo_id=0, rough_id=1,
                             seed=7, d_images=w)
    g_img = img.cpu().numpy()
    for v in range(V):
        _compare(g_img[v], o_img[v], f"adjoint forward v{v}")
    _cmp_grad(ga.cpu().numpy(), o_ga, "albedo gradient")
    _cmp_grad(gr.cpu().numpy(), o_gr, "roughness gradient")
    # the cached first pass and the recomputed one give the same gradients
    _cmp_grad(ga2.cpu().numpy(), ga.cpu().numpy(), "albedo gradient, recomputed first pass")
    _cmp_grad(gr2.cpu().numpy(), gr.cpu().numpy(), "roughness gradient, recomputed first pass")
    adj.close()


def test_adjoint_albedo_central_differences_on_gpu(gpu_ctx):
    """h = 1e-3 central differences of the GPU forward itself on 64 random
    texels (paths do not depend on albedo: the estimator is a polynomial in
    it, so the differences match to float rounding).  In float32 the
    difference of two images carries about 1e-7 of each pixel, i.e. 5e-5 of
    it after dividing by 2h, hence the loose bar here; the 1e-3 contract is
    checked against float64 central differences of the oracle's estimator
    (test_gpu_fullsize.py)."""
    print("[gpu] adjoint albedo central differences", flush=True)
    from lumen_amd.adjoint import Adjoint
    W = H = 32
    sc, alb, rough = _scene()
    rng = np.random.default_rng(4)
    w = torch.from_numpy(rng.normal(size=(1, H, W, 3)).astype(np.float32)).cuda()
    adj = Adjoint(sc, gpu_ctx, W, H, spp=4, variants=1, albedo_texture=0, roughness_texture=1, max_depth=5)
    ta, tr = torch.from_numpy(alb[None].copy()).cuda(), torch.from_numpy(rough[None].copy()).cuda()
    img, cache = adj.forward(ta, tr, seed=3)
    ga, _ = adj.backward(ta, tr, w, seed=3, cache=cache)
    ga = ga.cpu().numpy()
    flat = np.flatnonzero(np.abs(ga) > 1e-6 * np.abs(ga).max())
    picks = rng.choice(flat, size=64, replace=False)
    h = 1e-3
    for f in picks:
        idx = np.unravel_index(f, ga.shape)
        ap, am = ta.clone(), ta.clone()
        ap[idx] += h
        am[idx] -= h
        ip, _ = adj.forward(ap, tr, seed=3, cache=False)
        im, _ = adj.forward(am, tr, seed=3, cache=False)
        fd = float(((ip.double() - im.double()) * w.double()).sum()) / (2 * h)
        assert ga[idx] == pytest.approx(fd, rel=3e-2, abs=3e-3 * np.abs(ga).max()), idx
    adj.close()


def test_adjoint_counting_variants(gpu_ctx):
    """The counting kernels (lt_adjoint_set_counting) give the same images bit
    for bit and the same gradients up to the order of their float atomics,
    and count the work: at least one camera ray per
    sample, at least one node per ray; the cached replay reads the forward's
    path records and traces no closest-hit ray, only the occlusion rays the
    forward skipped (lights whose contribution is zero but whose derivative
    is not); a replay without the cache traces every closest-hit ray twice
    (the forward again, then the replay)."""
    print("[gpu] adjoint counting variants", flush=True)
    from lumen_amd.adjoint import Adjoint
    W, H, V, spp = 24, 16, 2, 2
    sc, alb, rough = _scene()
    albs, roughs = _variants(alb, rough, V, np.random.default_rng(3))
    w = torch.from_numpy(np.random.default_rng(4).normal(size=(V, H, W, 3)).astype(np.float32)).cuda()
    adj = Adjoint(sc, gpu_ctx, W, H, spp=spp, variants=V, albedo_texture=0, roughness_texture=1, max_depth=6)
    ta, tr = torch.from_numpy(albs).cuda(), torch.from_numpy(roughs).cuda()
    img, cache = adj.forward(ta, tr, seed=3)
    ga, gr = adj.backward(ta, tr, w, seed=3, cache=cache)
    adj.set_counting(True)
    img_c, cache_c = adj.forward(ta, tr, seed=3)
    cf = adj.counts()
    ga_c, gr_c = adj.backward(ta, tr, w, seed=3, cache=cache_c)
    cb = adj.counts()
    adj.backward(ta, tr, w, seed=3, cache=None)
    cn = adj.counts()
    adj.set_counting(False)
    torch.cuda.synchronize()
    # the radiance and the flag words are written for every sample; hit
    # records only for the bounces a path reaches
    n_rad = spp * 3 * V * W * H
    n_flags = spp * V * W * H
    assert torch.equal(img, img_c) and torch.equal(cache[:n_rad], cache_c[:n_rad])
    assert torch.equal(cache[-n_flags:], cache_c[-n_flags:])
    _cmp_grad(ga_c.cpu().numpy(), ga.cpu().numpy(), "albedo gradient, counting variant")
    _cmp_grad(gr_c.cpu().numpy(), gr.cpu().numpy(), "roughness gradient, counting variant")
    print(f"[gpu] adjoint counts forward {cf}, backward {cb}", flush=True)
    assert cf["closest"] >= V * W * H * spp and cf["nodes"] >= cf["closest"] + cf["shadow"]
    assert cf["tris"] > 0 and cf["shadow"] > 0 and cf["shadow_nodes"] <= cf["nodes"]
    assert cb["closest"] == 0 and cb["nodes"] == cb["shadow_nodes"] and cb["tris"] == cb["shadow_tris"]
    assert cb["shadow"] < cf["shadow"]
    assert cn["closest"] == 2 * cf["closest"] and cn["shadow"] >= cf["shadow"]
    assert cn["nodes"] - cn["shadow_nodes"] == 2 * (cf["nodes"] - cf["shadow_nodes"])
    adj.close()


def test_autograd_function_matches_direct_backward(gpu_ctx):
    print("[gpu] adjoint autograd", flush=True)
    from lumen_amd.adjoint import Adjoint, render_textures
    W, H, V = 32, 24, 2
    sc, alb, rough = _scene()
    rng = np.random.default_rng(9)
    albs, roughs = _variants(alb, rough, V, rng)
    adj = Adjoint(sc, gpu_ctx, W, H, spp=2, variants=V, albedo_texture=0, roughness_texture=1, max_depth=6)
    ta = torch.from_numpy(albs).cuda().requires_grad_(True)
    tr = torch.from_numpy(roughs).cuda().requires_grad_(True)
    target = torch.rand(V, H, W, 3, device="cuda")
    img = render_textures(adj, ta, tr, seed=5)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    d = (2 * (img.detach() - target)).contiguous()
    ga, gr = adj.backward(ta.detach(), tr.detach(), d, seed=5, cache=None)
    torch.cuda.synchronize()
    _cmp_grad(ta.grad.cpu().numpy(), ga.cpu().numpy(), "autograd albedo")
    _cmp_grad(tr.grad.cpu().numpy(), gr.cpu().numpy(), "autograd roughness")
    adj.close()


def test_adjoint_rejects_unsupported_materials(gpu_ctx):
    import ctypes as C
    from lumen_amd import _abi
    from lumen_amd.adjoint import Adjoint
    sc, _, _ = _scene()
    d = sc.desc
    wall = next(m for m in range(d.num_materials) if d.materials[m].texture_id == 0)
    d.materials[wall].metallic = 0.5
    gpu_ctx.upload(sc)
    gpu_ctx.build_accel()
    with pytest.raises(la.LtError, match="LT_ERR_UNSUPPORTED"):
        Adjoint(sc, gpu_ctx, 16, 16, spp=1, variants=1, albedo_texture=0, roughness_texture=1)
    d.materials[wall].metallic = 0.0
    cfg = _abi.lt_adjoint_config()
    cfg.width = cfg.height = 16
    cfg.max_depth, cfg.spp, cfg.variants = 4, 1, 1
    cfg.albedo_texture, cfg.roughness_texture = 7, -1
    h = C.c_void_p()
    assert la.lib().lt_adjoint_create(gpu_ctx.handle, C.byref(cfg), C.byref(h)) == 1  # LT_ERR_INVALID_ARG
