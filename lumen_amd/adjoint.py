"""Differentiable rendering of texture variants (lt_adjoint_*) as a
``torch.autograd.Function``.

No reference counterpart: the reference renders but does not differentiate.
The forward is the Path integrator (path.rgen:26-104) without roulette,
averaging ``spp`` samples per pixel; the backward is path-replay
backpropagation with detached sampling on the HIP kernels of
``csrc/adjoint.hip``.  Tensors stay on the GPU: the C ABI takes their device
pointers and torch's current stream, so the launches order with the
surrounding torch work.

    adj = Adjoint(scene, ctx, 512, 512, spp=32, variants=32,
                  albedo_texture=0, roughness_texture=1)
    img = render_textures(adj, albedo, roughness, seed)   # [V, H, W, 3]
    loss = ((img - target) ** 2).sum(); loss.backward()
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _abi
from ._abi import check, lib


class Adjoint:
    """One configured adjoint over the scene uploaded to ``ctx`` (uploads
    ``scene`` and builds its BVH when ``ctx`` holds another one)."""

    def __init__(self, scene, ctx, width: int, height: int, spp: int = 32, variants: int = 1,
                 albedo_texture: int = -1, roughness_texture: int = -1, max_depth: int = 6,
                 sky_col=(0.0, 0.0, 0.0), camera=None):
        if ctx.scene is not scene:
            ctx.upload(scene)
            ctx.build_accel()
        self.ctx, self.scene = ctx, scene
        self.width, self.height, self.spp, self.variants = width, height, spp, variants
        self.camera = camera if camera is not None else scene.camera(width, height)
        cfg = _abi.lt_adjoint_config()
        cfg.width, cfg.height, cfg.max_depth, cfg.spp, cfg.variants = width, height, max_depth, spp, variants
        cfg.albedo_texture, cfg.roughness_texture = albedo_texture, roughness_texture
        cfg.sky_col[:] = list(sky_col)
        self.cfg = cfg
        h = C.c_void_p()
        check(lib().lt_adjoint_create(ctx.handle, C.byref(cfg), C.byref(h)), ctx.handle, "lt_adjoint_create")
        self._h = h
        dims = [C.c_uint32() for _ in range(4)]
        check(lib().lt_adjoint_texture_size(h, *[C.byref(d) for d in dims]), ctx.handle, "lt_adjoint_texture_size")
        self.albedo_shape = (variants, dims[1].value, dims[0].value, 3)
        self.rough_shape = (variants, dims[3].value, dims[2].value)
        n = C.c_size_t()
        check(lib().lt_adjoint_cache_floats(h, C.byref(n)), ctx.handle, "lt_adjoint_cache_floats")
        self.cache_floats = n.value

    @property
    def image_shape(self):
        return (self.variants, self.height, self.width, 3)

    def _check(self, t, shape, name):
        if t is None:
            return None
        if tuple(t.shape) != tuple(shape) or t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
            raise ValueError(f"{name}: expected a contiguous float32 CUDA tensor of shape {tuple(shape)}, "
                             f"got {tuple(t.shape)} {t.dtype} {t.device}")
        return t

    @staticmethod
    def _ptr(t):
        return None if t is None else C.c_void_p(t.data_ptr())

    def forward(self, albedo, roughness, seed: int = 0, cache: bool = True):
        """images [V, H, W, 3]; with cache=True also the per-sample radiance
        the backward reuses (spp * 3 * items floats)."""
        a = self._check(albedo, self.albedo_shape, "albedo") if self.cfg.albedo_texture >= 0 else None
        r = self._check(roughness, self.rough_shape, "roughness") if self.cfg.roughness_texture >= 0 else None
        dev = (a if a is not None else r).device
        img = torch.empty(self.image_shape, dtype=torch.float32, device=dev)
        cache_t = torch.empty(self.cache_floats, dtype=torch.float32, device=dev) if cache else None
        stream = torch.cuda.current_stream(dev).cuda_stream
        check(lib().lt_adjoint_forward(self._h, C.byref(self.camera), seed, self._ptr(a), self._ptr(r),
                                       self._ptr(img), self._ptr(cache_t), C.c_void_p(stream)),
              self.ctx.handle, "lt_adjoint_forward")
        return img, cache_t

    def backward(self, albedo, roughness, d_images, seed: int = 0, cache=None):
        """(grad_albedo, grad_roughness) of sum(d_images * images)."""
        a = self._check(albedo, self.albedo_shape, "albedo") if self.cfg.albedo_texture >= 0 else None
        r = self._check(roughness, self.rough_shape, "roughness") if self.cfg.roughness_texture >= 0 else None
        d = self._check(d_images.contiguous(), self.image_shape, "d_images")
        if cache is not None:
            self._check(cache, (self.cache_floats,), "cache")
        ga = torch.empty_like(a) if a is not None else None
        gr = torch.empty_like(r) if r is not None else None
        stream = torch.cuda.current_stream(d.device).cuda_stream
        check(lib().lt_adjoint_backward(self._h, C.byref(self.camera), seed, self._ptr(a), self._ptr(r),
                                        self._ptr(d), self._ptr(cache), self._ptr(ga), self._ptr(gr),
                                        C.c_void_p(stream)),
              self.ctx.handle, "lt_adjoint_backward")
        return ga, gr

    def set_counting(self, on: bool) -> None:
        """Counting variants of the kernels for later calls (measurement)."""
        check(lib().lt_adjoint_set_counting(self._h, int(on)), self.ctx.handle, "lt_adjoint_set_counting")

    def counts(self) -> dict:
        """Work of the last forward / backward call made while counting."""
        c = (C.c_uint64 * 6)()
        check(lib().lt_adjoint_counts(self._h, c), self.ctx.handle, "lt_adjoint_counts")
        return dict(zip(("closest", "shadow", "nodes", "tris", "shadow_nodes", "shadow_tris"), c[:]))

    def close(self):
        if getattr(self, "_h", None):
            lib().lt_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _RenderTextures(torch.autograd.Function):
    @staticmethod
    def forward(ctx, albedo, roughness, adj: Adjoint, seed: int, use_cache: bool):
        img, cache = adj.forward(albedo, roughness, seed, cache=use_cache)
        ctx.adj, ctx.seed = adj, seed
        ctx.save_for_backward(albedo, roughness, cache)
        return img

    @staticmethod
    def backward(ctx, grad_img):
        albedo, roughness, cache = ctx.saved_tensors
        ga, gr = ctx.adj.backward(albedo, roughness, grad_img.contiguous(), ctx.seed, cache)
        return ga, gr, None, None, None


def render_textures(adj: Adjoint, albedo, roughness, seed: int = 0, use_cache: bool = True):
    """Differentiable render of the texture variants: [V, H, W, 3].

    ``use_cache`` keeps each sample's radiance from the forward (12 B per
    sample in HBM) and, for path lengths up to 10, its path records (32 B
    per bounce and 4 B of flags per sample), so the backward replays every
    path from its records instead of tracing it again."""
    if albedo is None:
        albedo = torch.zeros(0, device="cuda")
    if roughness is None:
        roughness = torch.zeros(0, device="cuda")
    return _RenderTextures.apply(albedo, roughness, adj, seed, use_cache)
