"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker / the timed CPU baseline; the product
(lumen_amd/) never does.  Build with ``make -C oracle``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# ORACLE_LIB selects another build of the same sources (e.g. _build/liboracle_fma.so)
LIB_PATH = Path(os.environ.get("ORACLE_LIB") or (HERE / "_build" / "liboracle.so"))


REF_TREE = Path("/root/reference")
REF_STBI = HERE / "_ref" / "libref_stbi.so"


def build() -> Path:
    r = subprocess.run(["make", "-C", str(HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    if (REF_TREE / "libs" / "stb_image" / "stb_image.h").exists():
        # the reference's own decoder, compiled from its tree (oracle/_ref)
        r = subprocess.run(["make", "-C", str(HERE), "ref", f"REF={REF_TREE}"], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"oracle/_ref build failed\n{r.stdout}\n{r.stderr}")
    return LIB_PATH


def _f3(v):
    return np.ascontiguousarray(np.asarray(v, np.float32).reshape(3))


def scattering(origin, direction, length, light_dir, light_col) -> np.ndarray:
    """integrate_scattering (atmosphere.glsl:143-205) of the oracle."""
    a = [_f3(origin), _f3(direction), _f3(light_dir), _f3(light_col)]
    out = np.zeros(3, np.float32)
    lib().ora_scattering(a[0].ctypes.data, a[1].ctypes.data, float(length), a[2].ctypes.data, a[3].ctypes.data,
                         out.ctypes.data)
    return out


def phase(which: str, costh: float, g: float = 0.85) -> float:
    return float(lib().ora_phase(0 if which == "rayleigh" else 1, costh, g))


def ref_stbi_load(path: str):
    """RGBA8 image decoded by the reference's stb_image (oracle/_ref), or
    None where that library is not built."""
    if not REF_STBI.exists():
        return None
    L = C.CDLL(str(REF_STBI))
    L.ref_stbi_load_rgba8.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    w, h = C.c_int(), C.c_int()
    if L.ref_stbi_load_rgba8(path.encode(), None, 0, C.byref(w), C.byref(h)) != 0:
        raise RuntimeError(f"stb_image cannot decode {path}")
    out = np.zeros((h.value, w.value, 4), np.uint8)
    if L.ref_stbi_load_rgba8(path.encode(), out.ctypes.data, out.size, C.byref(w), C.byref(h)) != 0:
        raise RuntimeError(f"stb_image cannot decode {path}")
    return out


class PathParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_depth", C.c_int32),
                ("sky_col", C.c_float * 3), ("direct_lighting", C.c_uint32), ("frame_start", C.c_uint32),
                ("frames", C.c_uint32), ("seed", C.c_uint32), ("x0", C.c_uint32), ("y0", C.c_uint32),
                ("x1", C.c_uint32), ("y1", C.c_uint32), ("nee_pick", C.c_uint32)]


class BdptParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_depth", C.c_int32),
                ("frame_start", C.c_uint32), ("frames", C.c_uint32), ("seed", C.c_uint32), ("power", C.c_uint32),
                ("x0", C.c_uint32), ("y0", C.c_uint32), ("x1", C.c_uint32), ("y1", C.c_uint32)]


class VcmParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_depth", C.c_int32),
                ("frame_start", C.c_uint32), ("frames", C.c_uint32), ("seed", C.c_uint32),
                ("radius_factor", C.c_float), ("alpha", C.c_float), ("use_vc", C.c_int32), ("use_vm", C.c_int32),
                ("corrected", C.c_uint32), ("sky_col", C.c_float * 3)]


class SppmParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_depth", C.c_int32), ("frame_start", C.c_uint32),
                ("frames", C.c_uint32), ("seed", C.c_uint32), ("base_radius", C.c_float), ("sky_col", C.c_float * 3)]


class AdjParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_depth", C.c_int32), ("spp", C.c_uint32),
                ("seed", C.c_uint32), ("sky_col", C.c_float * 3), ("albedo_id", C.c_int32), ("rough_id", C.c_int32),
                ("aw", C.c_uint32), ("ah", C.c_uint32), ("rw", C.c_uint32), ("rh", C.c_uint32)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        F = C.POINTER(C.c_float)
        L.ora_pcg4d.argtypes = [P, P, C.c_size_t]
        L.ora_uint_to_float.argtypes = [C.c_uint32]
        L.ora_uint_to_float.restype = C.c_float
        L.ora_rand_seq.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_size_t, P]
        L.ora_offset_ray.argtypes = [P, P, P, C.c_int]
        L.ora_concentric_disk.argtypes = [P, P]
        L.ora_onb.argtypes = [P, P, P]
        L.ora_sample_bsdf.argtypes = [P, P, P, C.c_uint32, C.c_int, P, P, P, F, F]
        L.ora_eval_bsdf.argtypes = [P, P, P, P, C.c_uint32, C.c_int, C.c_int, P, F, F]
        L.ora_bsdf_pdf.argtypes = [P, P, P, P, C.c_int]
        L.ora_bsdf_pdf.restype = C.c_float
        L.ora_fresnel_dielectric.argtypes = [C.c_float, C.c_float, C.c_int]
        L.ora_fresnel_dielectric.restype = C.c_float
        L.ora_fresnel_conductor.argtypes = [C.c_float, C.c_float, C.c_float]
        L.ora_fresnel_conductor.restype = C.c_float
        L.ora_D_ggx_iso.argtypes = [C.c_float, C.c_float]
        L.ora_D_ggx_iso.restype = C.c_float
        L.ora_sample_light_Li.argtypes = [P, P, P, P]
        L.ora_accel_build.argtypes = [P]
        L.ora_accel_build.restype = P
        L.ora_accel_free.argtypes = [P]
        L.ora_trace_closest.argtypes = [P, P, C.c_size_t, P, C.c_int, C.c_int]
        L.ora_trace_any.argtypes = [P, P, C.c_size_t, P, C.c_int, C.c_int]
        L.ora_payload.argtypes = [P, P, P, P]
        L.ora_atmosphere.argtypes = [P, P, P, P, C.c_float, P]
        L.ora_scattering.argtypes = [P, P, C.c_float, P, P, P]
        L.ora_phase.argtypes = [C.c_int, C.c_float, C.c_float]
        L.ora_phase.restype = C.c_float
        L.ora_render_path.argtypes = [P, P, P, C.POINTER(PathParams), P, C.c_int, C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_uint64)]
        L.ora_render_path.restype = C.c_double
        L.ora_render_bdpt.argtypes = [P, P, P, C.POINTER(BdptParams), P, P, C.c_int, C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_uint64)]
        L.ora_render_bdpt.restype = C.c_double
        L.ora_sample_light_Le.argtypes = [P, P, P]
        L.ora_render_vcm.argtypes = [P, P, P, C.POINTER(VcmParams), P, P, C.c_int, C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_uint64)]
        L.ora_render_vcm.restype = C.c_double
        L.ora_vcm_grid_selfcheck.argtypes = [P, P, P, C.POINTER(VcmParams), C.c_int]
        L.ora_vcm_words.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.ora_adjoint.argtypes = [P, P, P, C.POINTER(AdjParams), C.c_uint32, P, P, P, P, P, P, P, P, C.c_int]
        L.ora_adjoint.restype = C.c_double
        L.ora_bsdf_eval_d.argtypes = [P, P, P, P, C.c_int, P]
        L.ora_vcm_radius_grid.argtypes = [P, C.c_float, C.c_float, C.c_uint32, C.POINTER(C.c_float),
                                          C.POINTER(C.c_int32)]
        L.ora_hardware_threads.restype = C.c_int
        L.ora_sppm_create.argtypes = [C.c_uint32, C.c_uint32]
        L.ora_sppm_create.restype = P
        L.ora_sppm_free.argtypes = [P]
        L.ora_sppm_word.argtypes = [C.c_uint32, C.c_uint32]
        L.ora_sppm_word.restype = C.c_uint32
        L.ora_render_sppm.argtypes = [P, P, P, C.POINTER(SppmParams), P, P, P, C.c_int, C.POINTER(C.c_uint64)]
        L.ora_render_sppm.restype = C.c_double
        _lib = L
    return _lib


def _f3(v):
    return np.ascontiguousarray(v, np.float32)


def pcg4d(v: np.ndarray) -> np.ndarray:
    v = np.ascontiguousarray(v, np.uint32).reshape(-1, 4)
    out = np.zeros_like(v)
    lib().ora_pcg4d(v.ctypes.data, out.ctypes.data, len(v))
    return out


def rand_seq(x, y, frame, n, w0=0) -> np.ndarray:
    out = np.zeros(n, np.float32)
    lib().ora_rand_seq(x, y, frame, w0, n, out.ctypes.data)
    return out


def offset_ray(p, n, which=1) -> np.ndarray:
    p, n = _f3(p), _f3(n)
    out = np.zeros(3, np.float32)
    lib().ora_offset_ray(p.ctypes.data, n.ctypes.data, out.ctypes.data, which)
    return out


def concentric_disk(r) -> np.ndarray:
    r = _f3(r)
    out = np.zeros(2, np.float32)
    lib().ora_concentric_disk(r.ctypes.data, out.ctypes.data)
    return out


def onb(n):
    n = _f3(n)
    b1, b2 = np.zeros(3, np.float32), np.zeros(3, np.float32)
    lib().ora_onb(n.ctypes.data, b1.ctypes.data, b2.ctypes.data)
    return b1, b2


def sample_bsdf(mat, n_s, wo, rands, mode=1, side=True):
    n_s, wo, rands = _f3(n_s), _f3(wo), _f3(rands)
    wi, f = np.zeros(3, np.float32), np.zeros(3, np.float32)
    pdf, cos = C.c_float(), C.c_float()
    lib().ora_sample_bsdf(C.byref(mat), n_s.ctypes.data, wo.ctypes.data, mode, int(side), rands.ctypes.data,
                          wi.ctypes.data, f.ctypes.data, C.byref(pdf), C.byref(cos))
    return wi, f, pdf.value, cos.value


def eval_bsdf(mat, n_s, wo, wi, mode=1, side=True, rev=True):
    n_s, wo, wi = _f3(n_s), _f3(wo), _f3(wi)
    f = np.zeros(3, np.float32)
    pdf, pdf_rev = C.c_float(), C.c_float()
    lib().ora_eval_bsdf(C.byref(mat), n_s.ctypes.data, wo.ctypes.data, wi.ctypes.data, mode, int(side), int(rev),
                        f.ctypes.data, C.byref(pdf), C.byref(pdf_rev))
    return f, pdf.value, pdf_rev.value


def bsdf_eval_d(mat, n_s, wo, wi, side=True):
    """eval_bsdf and its albedo / roughness derivatives: (f, df/dalbedo, df/droughness)."""
    n_s, wo, wi = _f3(n_s), _f3(wo), _f3(wi)
    out = np.zeros(9, np.float32)
    lib().ora_bsdf_eval_d(C.byref(mat), n_s.ctypes.data, wo.ctypes.data, wi.ctypes.data, int(side), out.ctypes.data)
    return out[0:3], out[3:6], out[6:9]


def bsdf_pdf(mat, n_s, wo, wi, side=True) -> float:
    n_s, wo, wi = _f3(n_s), _f3(wo), _f3(wi)
    return lib().ora_bsdf_pdf(C.byref(mat), n_s.ctypes.data, wo.ctypes.data, wi.ctypes.data, int(side))


def sample_light_Li(desc, rands4, p):
    rands4, p = _f3(rands4), _f3(p)
    out = np.zeros(12, np.float32)
    lib().ora_sample_light_Li(C.byref(desc), rands4.ctypes.data, p.ctypes.data, out.ctypes.data)
    return out


def sample_light_Le(desc, rands6):
    """[Le3, pos3, wi3, n3, cos_from_light, pdf_pos_a, pdf_dir_w, flags]"""
    rands6 = _f3(rands6)
    out = np.zeros(16, np.float32)
    lib().ora_sample_light_Le(C.byref(desc), rands6.ctypes.data, out.ctypes.data)
    return out


def vcm_merge_stats():
    """(queries, queries with > 64 photons in the radius, largest count) of
    the last corrected-estimator render_vcm."""
    out = (C.c_uint64 * 3)()
    lib().ora_vcm_merge_stats(out)
    return tuple(out)


def vcm_words(seed: int, frame_num: int):
    lw, ew = C.c_uint32(), C.c_uint32()
    lib().ora_vcm_words(seed, frame_num, C.byref(lw), C.byref(ew))
    return lw.value, ew.value


def vcm_radius_grid(desc, radius_factor, alpha, frame_num):
    r = C.c_float()
    g = (C.c_int32 * 3)()
    lib().ora_vcm_radius_grid(C.byref(desc), radius_factor, alpha, frame_num, C.byref(r), g)
    return r.value, tuple(g)


class Accel:
    """Host BVH2 over the scene's triangles (+ brute force)."""

    def __init__(self, desc):
        self.desc = desc
        self._h = lib().ora_accel_build(C.byref(desc))

    def closest(self, rays: np.ndarray, brute=False, threads=0) -> np.ndarray:
        from lumen_amd._abi import HIT_DTYPE, RAY_DTYPE  # record layouts only
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        lib().ora_trace_closest(self._h, rays.ctypes.data, len(rays), hits.ctypes.data, int(brute),
                                threads or lib().ora_hardware_threads())
        return hits

    def any(self, rays: np.ndarray, brute=False, threads=0) -> np.ndarray:
        from lumen_amd._abi import RAY_DTYPE
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        occ = np.zeros(len(rays), np.uint8)
        lib().ora_trace_any(self._h, rays.ctypes.data, len(rays), occ.ctypes.data, int(brute),
                            threads or lib().ora_hardware_threads())
        return occ

    def atmosphere(self, sky_col, origin, direction, length) -> np.ndarray:
        """shade_atmosphere (commons.glsl:156-168) of this scene."""
        a = [_f3(sky_col), _f3(origin), _f3(direction)]
        out = np.zeros(3, np.float32)
        lib().ora_atmosphere(C.byref(self.desc), a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data, float(length),
                             out.ctypes.data)
        return out

    def payload(self, hit, ray) -> np.ndarray:
        out = np.zeros(13, np.float32)
        h = np.ascontiguousarray(hit)
        r = np.ascontiguousarray(ray)
        lib().ora_payload(C.byref(self.desc), h.ctypes.data, r.ctypes.data, out.ctypes.data)
        return out

    def render_path(self, camera, width, height, max_depth, sky_col=(0, 0, 0), direct_lighting=1, frame_start=0,
                    frames=1, seed=0, window=None, accum=None, threads=0, nee_pick=0):
        """Runs the reference Path integrator on the CPU; returns (image, seconds, closest, shadow).
        nee_pick=1: the NEE sample is weighted by its own pick probability."""
        p = PathParams()
        p.nee_pick = nee_pick
        p.width, p.height, p.max_depth = width, height, max_depth
        p.sky_col[:] = list(sky_col)
        p.direct_lighting = direct_lighting
        p.frame_start, p.frames, p.seed = frame_start, frames, seed
        if window:
            p.x0, p.y0, p.x1, p.y1 = window
        if accum is None:
            accum = np.zeros((height, width, 4), np.float32)
        cl, sh = C.c_uint64(), C.c_uint64()
        secs = lib().ora_render_path(C.byref(self.desc), self._h, C.byref(camera), C.byref(p), accum.ctypes.data,
                                     threads or lib().ora_hardware_threads(), C.byref(cl), C.byref(sh))
        return accum, secs, cl.value, sh.value

    def render_bdpt(self, camera, width, height, max_depth, frame_start=0, frames=1, seed=0, power=False,
                    window=None, accum=None, splat=None, threads=0):
        """Runs the reference BDPT integrator on the CPU; returns (image, seconds, closest, shadow)."""
        p = BdptParams()
        p.width, p.height, p.max_depth = width, height, max_depth
        p.frame_start, p.frames, p.seed, p.power = frame_start, frames, seed, int(power)
        if window:
            p.x0, p.y0, p.x1, p.y1 = window
        if accum is None:
            accum = np.zeros((height, width, 4), np.float32)
        if splat is None:
            splat = np.zeros((height, width, 3), np.float32)
        cl, sh = C.c_uint64(), C.c_uint64()
        secs = lib().ora_render_bdpt(C.byref(self.desc), self._h, C.byref(camera), C.byref(p), accum.ctypes.data,
                                     splat.ctypes.data, threads or lib().ora_hardware_threads(), C.byref(cl),
                                     C.byref(sh))
        return accum, secs, cl.value, sh.value

    def render_vcm(self, camera, width, height, max_depth, frame_start=0, frames=1, seed=0, radius_factor=0.025,
                   alpha=2.0 / 3.0, use_vc=True, use_vm=True, corrected=False, sky_col=(0, 0, 0), accum=None,
                   splat=None, threads=0):
        """Runs the reference VCM integrator (or its corrected form) on the CPU;
        returns (image, seconds, closest, shadow)."""
        p = VcmParams()
        p.width, p.height, p.max_depth = width, height, max_depth
        p.frame_start, p.frames, p.seed = frame_start, frames, seed
        p.radius_factor, p.alpha = radius_factor, alpha
        p.use_vc, p.use_vm, p.corrected = int(use_vc), int(use_vm), int(corrected)
        p.sky_col[:] = list(sky_col)
        if accum is None:
            accum = np.zeros((height, width, 4), np.float32)
        if splat is None:
            splat = np.zeros((height, width, 3), np.float32)
        cl, sh = C.c_uint64(), C.c_uint64()
        secs = lib().ora_render_vcm(C.byref(self.desc), self._h, C.byref(camera), C.byref(p), accum.ctypes.data,
                                    splat.ctypes.data, threads or lib().ora_hardware_threads(), C.byref(cl),
                                    C.byref(sh))
        return accum, secs, cl.value, sh.value

    def render_sppm(self, camera, width, height, max_depth, frame_start=0, frames=1, seed=0, base_radius=0.03,
                    sky_col=(0, 0, 0), state=None, accum=None, threads=0):
        """Runs the reference SPPM integrator on the CPU; returns (image,
        seconds, rays, grid [min3, max3, res3], state).  Pass the returned
        state back to continue at frame_start > 0."""
        p = SppmParams()
        p.width, p.height, p.max_depth = width, height, max_depth
        p.frame_start, p.frames, p.seed, p.base_radius = frame_start, frames, seed, base_radius
        p.sky_col[:] = list(sky_col)
        if state is None:
            state = SppmState(width, height)
        if accum is None:
            accum = np.zeros((height, width, 4), np.float32)
        grid = np.zeros(9, np.float32)
        rays = C.c_uint64()
        secs = lib().ora_render_sppm(C.byref(self.desc), self._h, C.byref(camera), C.byref(p), state.handle,
                                     accum.ctypes.data, grid.ctypes.data, threads or lib().ora_hardware_threads(),
                                     C.byref(rays))
        return accum, secs, rays.value, grid, state

    def vcm_grid_selfcheck(self, camera, width, height, max_depth, radius_factor, nq=200, seed=0):
        p = VcmParams()
        p.width, p.height, p.max_depth = width, height, max_depth
        p.seed, p.radius_factor, p.alpha = seed, radius_factor, 2.0 / 3.0
        return lib().ora_vcm_grid_selfcheck(C.byref(self.desc), self._h, C.byref(camera), C.byref(p), nq)

    def adjoint(self, camera, width, height, max_depth, spp, albedo=None, rough=None, albedo_id=-1, rough_id=-1,
                seed=0, sky_col=(0, 0, 0), d_images=None, eval_albedo=None, eval_rough=None, threads=0):
        """Differentiable Path estimator (no roulette) over texture variants.

        albedo: float32 [V, ah, aw, 3] replacing scene texture `albedo_id`;
        rough: float32 [V, rh, rw] replacing texture `rough_id`.  Without
        d_images returns images [V, H, W, 3]; with d_images (same shape)
        returns (grad_albedo, grad_rough) of sum(d_images * images) in float64.
        eval_albedo / eval_rough (forward only): BSDF values use these texels
        while sampling uses albedo / rough (detached-sampling estimator).
        """
        V = (albedo if albedo is not None else rough).shape[0]
        p = AdjParams()
        p.width, p.height, p.max_depth, p.spp, p.seed = width, height, max_depth, spp, seed
        p.sky_col[:] = list(sky_col)
        p.albedo_id, p.rough_id = albedo_id, rough_id
        a = np.ascontiguousarray(albedo if albedo is not None else np.zeros((V, 1, 1, 3)), np.float32)
        r = np.ascontiguousarray(rough if rough is not None else np.zeros((V, 1, 1)), np.float32)
        p.ah, p.aw = a.shape[1:3]
        p.rh, p.rw = r.shape[1:3]
        thr = threads or lib().ora_hardware_threads()
        if d_images is None:
            img = np.zeros((V, height, width, 3), np.float32)
            ea = None if eval_albedo is None else np.ascontiguousarray(eval_albedo, np.float32)
            er = None if eval_rough is None else np.ascontiguousarray(eval_rough, np.float32)
            lib().ora_adjoint(C.byref(self.desc), self._h, C.byref(camera), C.byref(p), V, a.ctypes.data,
                              r.ctypes.data, None if ea is None else ea.ctypes.data,
                              None if er is None else er.ctypes.data, img.ctypes.data, None, None, None, thr)
            return img
        d = np.ascontiguousarray(d_images, np.float32)
        ga = np.zeros(a.shape, np.float64)
        gr = np.zeros(r.shape, np.float64)
        lib().ora_adjoint(C.byref(self.desc), self._h, C.byref(camera), C.byref(p), V, a.ctypes.data,
                          r.ctypes.data, None, None, None, d.ctypes.data, ga.ctypes.data, gr.ctypes.data, thr)
        return ga, gr

    def __del__(self):
        try:
            lib().ora_accel_free(self._h)
        except Exception:
            pass


class AdjParams64(C.Structure):
    """ora_adj_params of the float64 build (its sky_col holds doubles)."""
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_depth", C.c_int32), ("spp", C.c_uint32),
                ("seed", C.c_uint32), ("sky_col", C.c_double * 3), ("albedo_id", C.c_int32), ("rough_id", C.c_int32),
                ("aw", C.c_uint32), ("ah", C.c_uint32), ("rw", C.c_uint32), ("rh", C.c_uint32)]


_lib64 = None


def lib64() -> C.CDLL:
    """The differentiable estimator compiled in float64 (oracle_f64.cpp)."""
    global _lib64
    if _lib64 is None:
        path = HERE / "_build" / "liboracle_f64.so"
        if not path.exists():
            build()
        L = C.CDLL(str(path))
        P = C.c_void_p
        L.ora64_accel_build.argtypes = [P]
        L.ora64_accel_build.restype = P
        L.ora64_accel_free.argtypes = [P]
        L.ora_adjoint.argtypes = [P, P, P, C.POINTER(AdjParams64), C.c_uint32, P, P, P, P, P, P, P, P, C.c_int]
        L.ora_adjoint.restype = C.c_double
        _lib64 = L
    return _lib64


class Accel64:
    """The forward of Accel.adjoint evaluated in float64: same scene, camera,
    seed, random numbers and sampling decisions (up to ties within a rounding
    error), so central differences of it carry float64 rounding only."""

    def __init__(self, desc):
        self.desc = desc
        self._h = lib64().ora64_accel_build(C.byref(desc))

    def adjoint(self, camera, width, height, max_depth, spp, albedo=None, rough=None, albedo_id=-1, rough_id=-1,
                seed=0, sky_col=(0, 0, 0), eval_albedo=None, eval_rough=None, threads=0):
        """images float64 [V, H, W, 3] of the texel variants (as Accel.adjoint's forward)."""
        V = (albedo if albedo is not None else rough).shape[0]
        p = AdjParams64()
        p.width, p.height, p.max_depth, p.spp, p.seed = width, height, max_depth, spp, seed
        p.sky_col[:] = list(sky_col)
        p.albedo_id, p.rough_id = albedo_id, rough_id
        a = np.ascontiguousarray(albedo if albedo is not None else np.zeros((V, 1, 1, 3)), np.float64)
        r = np.ascontiguousarray(rough if rough is not None else np.zeros((V, 1, 1)), np.float64)
        p.ah, p.aw = a.shape[1:3]
        p.rh, p.rw = r.shape[1:3]
        ea = None if eval_albedo is None else np.ascontiguousarray(eval_albedo, np.float64)
        er = None if eval_rough is None else np.ascontiguousarray(eval_rough, np.float64)
        img = np.zeros((V, height, width, 3), np.float64)
        lib64().ora_adjoint(C.byref(self.desc), self._h, C.byref(camera), C.byref(p), V, a.ctypes.data,
                            r.ctypes.data, None if ea is None else ea.ctypes.data,
                            None if er is None else er.ctypes.data, img.ctypes.data, None, None, None,
                            threads or lib().ora_hardware_threads())
        return img

    def __del__(self):
        try:
            lib64().ora64_accel_free(self._h)
        except Exception:
            pass


class SppmState:
    """Per-pixel SPPM data carried between frames (sppm_commons.h:27-42)."""

    def __init__(self, width, height):
        self.handle = lib().ora_sppm_create(width, height)

    def __del__(self):
        try:
            lib().ora_sppm_free(self.handle)
        except Exception:
            pass


def sppm_word(seed: int, frame_num: int) -> int:
    return lib().ora_sppm_word(seed, frame_num)


def bvh8_trace(nodes: np.ndarray, tris: np.ndarray, rays: np.ndarray, any_hit: bool = False):
    """Host traversal of an exported device BVH8: (hits or occluded, node counts, tri counts)."""
    from lumen_amd._abi import HIT_DTYPE, RAY_DTYPE  # record layouts only
    L = lib()
    P = C.c_void_p
    L.ora_bvh8_trace.argtypes = [P, C.c_uint32, P, P, C.c_size_t, P, P, C.c_int, P, P]
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    n = len(rays)
    hits = np.zeros(n, HIT_DTYPE)
    occ = np.zeros(n, np.uint8)
    nc = np.zeros(n, np.uint64)
    tc = np.zeros(n, np.uint64)
    nodes = np.ascontiguousarray(nodes)
    tris = np.ascontiguousarray(tris)
    L.ora_bvh8_trace(nodes.ctypes.data, len(nodes), tris.ctypes.data, rays.ctypes.data, n, hits.ctypes.data,
                     occ.ctypes.data, int(any_hit), nc.ctypes.data, tc.ctypes.data)
    return (occ if any_hit else hits), nc, tc


def hardware_threads() -> int:
    return lib().ora_hardware_threads()
