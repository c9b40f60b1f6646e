"""lumen_amd — MI355X-native light-transport core behind Lumen's Integrator API.

Python host mirror of the reference's plugin interface
(src/RayTracer/Integrator.h:14-33, RayTracer::create_integrator
src/RayTracer/RayTracer.cpp:244-282) over the C ABI of include/lt.h.  The
compute path is the HIP library lumen_amd/liblt.so; importing this package
without it raises (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import (HIT_DTYPE, LT_INTEGRATOR_BDPT, LT_INTEGRATOR_PATH, LT_INTEGRATOR_SPPM, LT_INTEGRATOR_VCM,
                   LT_PATH_MEGAKERNEL, LT_PATH_PERSISTENT, LT_PATH_WAVEFRONT, LT_PTR_DEVICE, LT_SCENE_CAUSTICS,
                   LT_SCENE_CORNELL, LT_SCENE_DIFF_WALL, LT_SCENE_TERRAIN, RAY_DTYPE, LtError, check, lib)

__all__ = ["Scene", "Context", "Integrator", "Path", "BDPT", "VCM", "SPPM", "create_integrator", "RAY_DTYPE", "HIT_DTYPE", "LtError",
           "lib", "make_rays", "save_exr", "load_exr"]

_KINDS = {"cornell": LT_SCENE_CORNELL, "terrain": LT_SCENE_TERRAIN, "caustics": LT_SCENE_CAUSTICS,
          "diff_wall": LT_SCENE_DIFF_WALL}


class Scene:
    """A host scene owned by liblt: the flattened LumenScene vectors
    (src/RayTracer/LumenScene.h:47-99) plus camera and integrator block."""

    def __init__(self, handle):
        self._h = handle
        self.desc = _abi.lt_scene_desc()
        check(lib().lt_scene_view(self._h, C.byref(self.desc)), what="lt_scene_view")

    @classmethod
    def procedural(cls, kind: str, seed: int = 0, **kw) -> "Scene":
        p = _abi.lt_procedural_params()
        p.seed = seed
        for k, v in kw.items():
            setattr(p, k, v)
        h = C.c_void_p()
        check(lib().lt_scene_procedural(_KINDS[kind], C.byref(p), C.byref(h)), what=f"procedural {kind}")
        return cls(h)

    @classmethod
    def load_json(cls, path: str) -> "Scene":
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        st = lib().lt_scene_load_json(str(path).encode(), C.byref(h), err, 512)
        check(st, what=f"lt_scene_load_json({path}): {err.value.decode()}")
        return cls(h)

    def camera(self, width: int, height: int) -> _abi.lt_camera:
        cam = _abi.lt_camera()
        check(lib().lt_scene_camera(self._h, width, height, C.byref(cam)), what="lt_scene_camera")
        return cam

    def set_camera(self, fov: float, pos, direction) -> None:
        p = (C.c_float * 3)(*pos)
        d = (C.c_float * 3)(*direction)
        check(lib().lt_scene_set_camera(self._h, fov, p, d), what="lt_scene_set_camera")

    def integrator_cfg(self) -> _abi.lt_scene_integrator_cfg:
        cfg = _abi.lt_scene_integrator_cfg()
        check(lib().lt_scene_integrator(self._h, C.byref(cfg)), what="lt_scene_integrator")
        return cfg

    def _mesh_ranges(self):
        """(index_offset, index_count, vertex_offset) per prim mesh."""
        d = self.desc
        pinfo = np.ctypeslib.as_array(C.cast(d.prim_infos, C.POINTER(C.c_uint32)), shape=(d.num_prim_infos, 12))
        if d.index_counts:
            counts = np.ctypeslib.as_array(d.index_counts, shape=(d.num_prim_infos,)).astype(np.int64)
        else:
            offs = list(pinfo[:, 0].astype(np.int64)) + [d.num_indices]
            counts = np.diff(offs)
        return pinfo[:, 0].astype(np.int64), counts, pinfo[:, 1].astype(np.int64)

    @property
    def num_triangles(self) -> int:
        """Triangles over every mesh, each instance of a shared geometry counted."""
        return int(self._mesh_ranges()[1].sum() // 3)

    def triangles(self) -> np.ndarray:
        """World-space vertex positions per triangle (mesh by mesh, instances
        through their world matrices), shape (T, 3, 3)."""
        d = self.desc
        vtx = np.ctypeslib.as_array(C.cast(d.vertices, C.POINTER(C.c_float)), shape=(d.num_vertices, 8))
        idx = np.ctypeslib.as_array(d.indices, shape=(d.num_indices,)).astype(np.int64)
        world = (np.ctypeslib.as_array(d.world_matrices, shape=(d.num_prim_infos, 16)) if d.world_matrices
                 else None)
        out = []
        for m, (off, cnt, vo) in enumerate(zip(*self._mesh_ranges())):
            p = vtx[idx[off:off + cnt] + vo, :3]
            if world is not None:
                W = world[m].reshape(4, 4).T.astype(np.float64)  # column-major
                p = (p.astype(np.float64) @ W[:3, :3].T + W[:3, 3]).astype(np.float32)
            out.append(p.reshape(-1, 3, 3))
        return np.concatenate(out) if out else np.zeros((0, 3, 3), np.float32)

    def close(self):
        if self._h:
            lib().lt_scene_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One HIP device and stream (lt_context)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        check(lib().lt_context_create(device, C.byref(self._h)), what="lt_context_create")
        self.scene = None

    @property
    def handle(self):
        return self._h

    def device_info(self):
        n = C.c_int()
        arch = C.create_string_buffer(64)
        check(lib().lt_device_info(self._h, C.byref(n), arch, 64), self._h, "lt_device_info")
        return n.value, arch.value.decode()

    def set_tuning(self, **kw) -> None:
        t = _abi.lt_tuning()
        for k, v in kw.items():
            setattr(t, k, int(v))
        check(lib().lt_set_tuning(self._h, C.byref(t)), self._h, "lt_set_tuning")

    def get_tuning(self) -> dict:
        t = _abi.lt_tuning()
        check(lib().lt_get_tuning(self._h, C.byref(t)), self._h, "lt_get_tuning")
        return {"path_blocks_per_cu": t.path_blocks_per_cu, "trace_blocks_per_cu": t.trace_blocks_per_cu,
                "refill_min": t.refill_min, "bdpt_blocks_per_cu": t.bdpt_blocks_per_cu, "regen_min": t.regen_min,
                "bdpt_chunk_tiles": t.bdpt_chunk_tiles, "wf_sort": t.wf_sort, "xcd_bands": t.xcd_bands}

    def upload(self, scene: Scene) -> None:
        check(lib().lt_scene_upload(self._h, C.byref(scene.desc)), self._h, "lt_scene_upload")
        self.scene = scene

    def build_accel(self, max_leaf_tris: int = 1, collapse: str = "sah", tree: str = "ploc") -> _abi.lt_build_stats:
        """Integrator::create_accel: the binary tree (PLOC clustering or the
        Morton-code LBVH) collapsed into BVH8 nodes (SAH-optimal or greedy)."""
        o = _abi.lt_build_opts()
        o.max_leaf_tris = max_leaf_tris
        o.collapse = {"sah": _abi.LT_COLLAPSE_SAH, "greedy": _abi.LT_COLLAPSE_GREEDY}[collapse]
        o.binary_tree = {"ploc": _abi.LT_TREE_PLOC, "lbvh": _abi.LT_TREE_LBVH}[tree]
        s = _abi.lt_build_stats()
        check(lib().lt_accel_build(self._h, C.byref(o), C.byref(s)), self._h, "lt_accel_build")
        return s

    def export_bvh(self, stats: _abi.lt_build_stats):
        """(nodes uint8[N,128], tris uint8[T,48]) of the built BVH8."""
        nodes = np.zeros((stats.num_nodes, 128), np.uint8)
        tris = np.zeros((stats.num_tris, 48), np.uint8)
        check(lib().lt_accel_export(self._h, nodes.ctypes.data, nodes.nbytes, tris.ctypes.data, tris.nbytes),
              self._h, "lt_accel_export")
        return nodes, tris

    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        check(lib().lt_trace_closest(self._h, rays.ctypes.data, len(rays), hits.ctypes.data, 0), self._h,
              "lt_trace_closest")
        return hits

    def trace_any(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        occ = np.zeros(len(rays), np.uint8)
        check(lib().lt_trace_any(self._h, rays.ctypes.data, len(rays), occ.ctypes.data, 0), self._h, "lt_trace_any")
        return occ

    def trace_count(self, rays: np.ndarray, any_hit: bool = False, device_ptr: int | None = None,
                    n: int | None = None):
        nodes, tris = C.c_uint64(), C.c_uint64()
        if device_ptr is not None:
            st = lib().lt_trace_count(self._h, device_ptr, n, int(any_hit), LT_PTR_DEVICE, C.byref(nodes),
                                      C.byref(tris))
        else:
            rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
            st = lib().lt_trace_count(self._h, rays.ctypes.data, len(rays), int(any_hit), 0, C.byref(nodes),
                                      C.byref(tris))
        check(st, self._h, "lt_trace_count")
        return nodes.value, tris.value

    # device-resident buffers
    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        check(lib().lt_device_alloc(self._h, nbytes, C.byref(p)), self._h, "lt_device_alloc")
        return p.value

    def free(self, ptr: int) -> None:
        check(lib().lt_device_free(self._h, ptr), self._h, "lt_device_free")

    def h2d(self, dst: int, arr: np.ndarray) -> None:
        arr = np.ascontiguousarray(arr)
        check(lib().lt_memcpy_h2d(self._h, dst, arr.ctypes.data, arr.nbytes), self._h, "lt_memcpy_h2d")

    def d2h(self, arr: np.ndarray, src: int) -> None:
        check(lib().lt_memcpy_d2h(self._h, arr.ctypes.data, src, arr.nbytes), self._h, "lt_memcpy_d2h")

    def trace_closest_device(self, d_rays: int, n: int, d_hits: int) -> None:
        check(lib().lt_trace_closest(self._h, d_rays, n, d_hits, LT_PTR_DEVICE), self._h, "lt_trace_closest")

    def trace_any_device(self, d_rays: int, n: int, d_occ: int) -> None:
        check(lib().lt_trace_any(self._h, d_rays, n, d_occ, LT_PTR_DEVICE), self._h, "lt_trace_any")

    def last_kernel_ms(self) -> float:
        ms = C.c_float()
        check(lib().lt_last_kernel_ms(self._h, C.byref(ms)), self._h, "lt_last_kernel_ms")
        return ms.value

    def stream_copy(self, d_dst: int, d_src: int, nbytes: int) -> float:
        """Device-to-device copy kernel; returns its duration in ms."""
        ms = C.c_float()
        check(lib().lt_stream_copy(self._h, d_dst, d_src, nbytes, C.byref(ms)), self._h, "lt_stream_copy")
        return ms.value

    def synchronize(self) -> None:
        check(lib().lt_synchronize(self._h), self._h, "lt_synchronize")

    def sq_error(self, a: np.ndarray, b: np.ndarray) -> float:
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        assert a.shape == b.shape and a.shape[-1] == 4
        out = C.c_double()
        n = a.size // 4
        check(lib().lt_image_sq_error(self._h, a.ctypes.data, b.ctypes.data, n, 0, C.byref(out)), self._h,
              "lt_image_sq_error")
        return out.value

    def rmse(self, a: np.ndarray, b: np.ndarray):
        """(Lumen's RMSE, standard RMSE) of two (H, W, 4) images over rgb:
        sqrt(S) / (3 N) as output_rmse.comp:23 computes it, and sqrt(S / (3 N)),
        with S = sum over pixels of |a.rgb - b.rgb|^2 reduced on the GPU."""
        S = self.sq_error(a, b)
        n = a.size // 4
        return (float(np.sqrt(S)) / (3 * n), float(np.sqrt(S / (3 * n)))) if n else (0.0, 0.0)

    def close(self):
        if self._h:
            lib().lt_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Integrator:
    """Mirror of Lumen's abstract Integrator (src/RayTracer/Integrator.h:14-33):
    init / render / update / destroy / create_accel, plus the public
    `output_tex`, `updated` and `frame_num` members."""

    kind = -1

    def __init__(self, lumen_scene: Scene, ctx: Context, width: int = 1920, height: int = 1080):
        self.lumen_scene = lumen_scene
        self.ctx = ctx
        self.width, self.height = width, height
        self.updated = False
        self.frame_num = 0
        self.seed = 0
        self._it = None
        self.camera = lumen_scene.camera(width, height)

    def _config(self):
        raise NotImplementedError

    def init(self) -> None:
        """Integrator::init (Integrator.cpp:8-50): uploads the scene and creates
        the RGBA32F output."""
        self.ctx.upload(self.lumen_scene)
        cfg = self._config()
        h = C.c_void_p()
        check(lib().lt_integrator_create(self.ctx.handle, self.kind, C.byref(cfg), C.sizeof(cfg), C.byref(h)),
              self.ctx.handle, "lt_integrator_create")
        self._it = h
        self.frame_num = 0

    def create_accel(self) -> _abi.lt_build_stats:
        """Integrator::create_accel (Integrator.cpp:137-160)."""
        return self.ctx.build_accel()

    def render(self, spp: int = 1) -> None:
        """Integrator::render: records (here: launches) one frame's work."""
        p = _abi.lt_frame_params()
        p.camera = self.camera
        p.frame_num = self.frame_num
        p.seed = self.seed
        p.spp = spp
        check(lib().lt_render(self._it, C.byref(p)), self.ctx.handle, "lt_render")
        self._pending = spp

    def update(self) -> bool:
        """Path::update (Path.cpp:61-68): advances frame_num, resets it on
        `updated`."""
        self.frame_num += getattr(self, "_pending", 1)
        self._pending = 1
        if self.updated:
            self.frame_num = 0
            self.updated = False
            return True
        return False

    def set_graph(self, on: bool) -> None:
        """Replay each frame from a captured hipGraph (Path, BDPT)."""
        check(lib().lt_render_set_graph(self._it, int(on)), self.ctx.handle, "lt_render_set_graph")

    def set_counting(self, on: bool) -> None:
        check(lib().lt_render_set_counting(self._it, int(on)), self.ctx.handle, "lt_render_set_counting")

    def set_timing(self, on: bool) -> None:
        """Per-phase device timing of later one-frame renders (stats().phase_ms)."""
        check(lib().lt_render_set_timing(self._it, int(on)), self.ctx.handle, "lt_render_set_timing")

    def stats(self) -> _abi.lt_render_stats:
        s = _abi.lt_render_stats()
        check(lib().lt_render_stats_get(self._it, C.byref(s)), self.ctx.handle, "lt_render_stats_get")
        return s

    def reset(self) -> None:
        check(lib().lt_accum_reset(self._it), self.ctx.handle, "lt_accum_reset")
        self.frame_num = 0

    @property
    def output_tex(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), np.float32)
        check(lib().lt_readback_rgba32f(self._it, out.ctypes.data, out.size), self.ctx.handle,
              "lt_readback_rgba32f")
        return out

    def destroy(self) -> None:
        if self._it:
            lib().lt_integrator_destroy(self._it)
            self._it = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Path(Integrator):
    """The unidirectional MIS path tracer (src/RayTracer/Path.{h,cpp})."""

    kind = LT_INTEGRATOR_PATH

    def __init__(self, lumen_scene, ctx, width=1920, height=1080, path_length=None, sky_col=None,
                 direct_lighting=True, strategy="megakernel", nee_pick="reference"):
        super().__init__(lumen_scene, ctx, width, height)
        self.nee_pick = {"reference": _abi.LT_NEE_REFERENCE, "corrected": _abi.LT_NEE_CORRECTED}[nee_pick]
        icfg = lumen_scene.integrator_cfg()
        self.path_length = path_length if path_length is not None else (icfg.path_length or 6)
        self.sky_col = sky_col if sky_col is not None else tuple(icfg.sky_col)
        self.direct_lighting = direct_lighting
        self.strategy = {"megakernel": LT_PATH_MEGAKERNEL, "wavefront": LT_PATH_WAVEFRONT,
                         "persistent": LT_PATH_PERSISTENT}[strategy]

    def _config(self):
        c = _abi.lt_path_config()
        c.width, c.height = self.width, self.height
        c.max_depth = self.path_length
        c.sky_col[:] = list(self.sky_col)
        c.direct_lighting = int(self.direct_lighting)
        c.strategy = self.strategy
        c.nee_pick = self.nee_pick
        return c


class BDPT(Integrator):
    """Bidirectional path tracing (src/RayTracer/BDPT.{h,cpp}): light and camera
    subpaths of up to path_length + 1 vertices, every (s, t) connection,
    light-tracing splats, MIS with the reference's balance heuristic
    (mis="balance") or the power heuristic (mis="power")."""

    kind = LT_INTEGRATOR_BDPT

    def __init__(self, lumen_scene, ctx, width=1920, height=1080, path_length=None, mis="balance"):
        super().__init__(lumen_scene, ctx, width, height)
        icfg = lumen_scene.integrator_cfg()
        self.path_length = path_length if path_length is not None else (icfg.path_length or 6)
        self.mis = {"balance": _abi.LT_MIS_BALANCE, "power": _abi.LT_MIS_POWER}[mis]

    def _config(self):
        c = _abi.lt_bdpt_config()
        c.width, c.height = self.width, self.height
        c.max_depth = self.path_length
        c.mis = self.mis
        return c


class VCM(Integrator):
    """Vertex connection and merging (src/RayTracer/VCM.{h,cpp}): one light
    subpath per pixel per frame, camera splats, NEE, connections to one
    random light subpath, and (enable_vm) photon merging over a hash grid
    whose radius shrinks as r_0 / (n+1)^((1-alpha)/2).  estimator="reference"
    reproduces the reference shader; "corrected" sums every connection and
    merge."""

    kind = LT_INTEGRATOR_VCM

    def __init__(self, lumen_scene, ctx, width=1920, height=1080, path_length=None, sky_col=None,
                 radius_factor=None, alpha=2.0 / 3.0, enable_vm=None, estimator="reference"):
        super().__init__(lumen_scene, ctx, width, height)
        icfg = lumen_scene.integrator_cfg()
        self.path_length = path_length if path_length is not None else (icfg.path_length or 6)
        self.sky_col = sky_col if sky_col is not None else tuple(icfg.sky_col)
        self.radius_factor = radius_factor if radius_factor is not None else (icfg.radius_factor or 0.025)
        self.alpha = alpha
        self.enable_vm = bool(icfg.enable_vm) if enable_vm is None else bool(enable_vm)
        self.estimator = {"reference": _abi.LT_VCM_REFERENCE, "corrected": _abi.LT_VCM_CORRECTED}[estimator]

    def _config(self):
        c = _abi.lt_vcm_config()
        c.width, c.height = self.width, self.height
        c.max_depth = self.path_length
        c.sky_col[:] = list(self.sky_col)
        c.radius_factor = self.radius_factor
        c.alpha = self.alpha
        c.enable_vm = int(self.enable_vm)
        c.estimator = self.estimator
        return c


class SPPM(Integrator):
    """Stochastic progressive photon mapping (src/RayTracer/SPPM.{h,cpp}):
    per frame an eye pass through pixel centres that records the first
    diffuse surface, a light pass whose photons fill a one-photon-per-bucket
    hash grid sized by the largest gather radius, and a gather with the
    progressive radius update (alpha = 2/3).  The per-pixel state persists
    from frame 0 on."""

    kind = LT_INTEGRATOR_SPPM

    def __init__(self, lumen_scene, ctx, width=1920, height=1080, path_length=None, sky_col=None,
                 base_radius=None):
        super().__init__(lumen_scene, ctx, width, height)
        icfg = lumen_scene.integrator_cfg()
        self.path_length = path_length if path_length is not None else (icfg.path_length or 6)
        self.sky_col = sky_col if sky_col is not None else tuple(icfg.sky_col)
        self.base_radius = base_radius if base_radius is not None else (icfg.base_radius or 0.03)

    def _config(self):
        c = _abi.lt_sppm_config()
        c.width, c.height = self.width, self.height
        c.max_depth = self.path_length
        c.sky_col[:] = list(self.sky_col)
        c.base_radius = self.base_radius
        return c


def create_integrator(name: str, lumen_scene: Scene, ctx: Context, width: int = 1920, height: int = 1080,
                      **kw) -> Integrator:
    """RayTracer::create_integrator (RayTracer.cpp:244-282), keyed by the
    IntegratorType names of SceneConfig.h:13 that this core builds."""
    n = name.lower()
    if n == "path":
        return Path(lumen_scene, ctx, width, height, **kw)
    if n == "bdpt":
        return BDPT(lumen_scene, ctx, width, height, **kw)
    if n == "vcm":
        return VCM(lumen_scene, ctx, width, height, **kw)
    if n == "sppm":
        return SPPM(lumen_scene, ctx, width, height, **kw)
    raise LtError(6, f"integrator '{name}' is not built")


def make_rays(origins: np.ndarray, dirs: np.ndarray, tmin=0.0, tmax=1e30) -> np.ndarray:
    r = np.zeros(len(origins), RAY_DTYPE)
    r["org"] = origins
    r["dir"] = dirs
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def save_exr(path, rgba: np.ndarray, half: bool = True) -> None:
    """Writes the RGB of an (H, W, 4) float32 image (row 0 at the top) as an
    OpenEXR file, channels B, G, R, stored as HALF (the reference's
    ImageUtils::save_exr, src/Framework/ImageUtils.cpp:23-89) or FLOAT."""
    img = np.ascontiguousarray(rgba, dtype=np.float32)
    if img.ndim != 3 or img.shape[2] != 4:
        raise ValueError(f"save_exr wants an (H, W, 4) image, got {img.shape}")
    err = C.create_string_buffer(512)
    st = lib().lt_exr_save(str(path).encode(), img.ctypes.data, img.shape[1], img.shape[0], 1 if half else 2, err, 512)
    check(st, what=f"lt_exr_save({path}): {err.value.decode()}")


def load_exr(path) -> np.ndarray:
    """Reads an OpenEXR file (NONE / ZIPS / ZIP scanlines) as an (H, W, 4)
    float32 image (ImageUtils::load_exr, src/Framework/ImageUtils.cpp:9-21)."""
    w, h = C.c_uint32(), C.c_uint32()
    err = C.create_string_buffer(512)
    st = lib().lt_exr_load(str(path).encode(), None, 0, C.byref(w), C.byref(h), err, 512)
    check(st, what=f"lt_exr_load({path}): {err.value.decode()}")
    img = np.empty((h.value, w.value, 4), np.float32)
    st = lib().lt_exr_load(str(path).encode(), img.ctypes.data, img.size, C.byref(w), C.byref(h), err, 512)
    check(st, what=f"lt_exr_load({path}): {err.value.decode()}")
    return img


def load_image_rgba8(path) -> np.ndarray:
    """Decodes a baseline JPEG texture as an (H, W, 4) uint8 image, row 0 at
    the top, alpha 255 (the scene loader's texture decode; the reference's
    stbi_load(..., 4), src/RayTracer/LumenScene.cpp:205-212)."""
    w, h = C.c_uint32(), C.c_uint32()
    err = C.create_string_buffer(512)
    st = lib().lt_image_load_rgba8(str(path).encode(), None, 0, C.byref(w), C.byref(h), err, 512)
    check(st, what=f"lt_image_load_rgba8({path}): {err.value.decode()}")
    img = np.empty((h.value, w.value, 4), np.uint8)
    st = lib().lt_image_load_rgba8(str(path).encode(), img.ctypes.data, img.size, C.byref(w), C.byref(h), err, 512)
    check(st, what=f"lt_image_load_rgba8({path}): {err.value.decode()}")
    return img
