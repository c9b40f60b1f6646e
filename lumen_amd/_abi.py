"""ctypes mirror of include/lt.h (structures, enums and function prototypes).

The shared library is loaded from lumen_amd/liblt.so; there is no fallback:
if it is missing the import fails with an error telling how to build it.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

# the built library (LT_LIBRARY names a variant build of the same sources,
# as the autotuner's build-time sweep makes)
LIB_PATH = Path(os.environ.get("LT_LIBRARY") or (Path(__file__).resolve().parent / "liblt.so"))

# ---- status codes -----------------------------------------------------------
LT_OK = 0
STATUS_NAMES = {
    0: "LT_OK", 1: "LT_ERR_INVALID_ARG", 2: "LT_ERR_HIP", 3: "LT_ERR_OUT_OF_MEMORY", 4: "LT_ERR_NOT_READY",
    5: "LT_ERR_PARSE", 6: "LT_ERR_UNSUPPORTED", 7: "LT_ERR_NO_DEVICE",
}

LT_BSDF_DIFFUSE, LT_BSDF_MIRROR, LT_BSDF_GLASS, LT_BSDF_DIELECTRIC, LT_BSDF_CONDUCTOR, LT_BSDF_PRINCIPLED = (
    1, 2, 4, 8, 16, 32)
LT_BSDF_FLAG_DIFFUSE, LT_BSDF_FLAG_SPECULAR, LT_BSDF_FLAG_GLOSSY, LT_BSDF_FLAG_REFLECTION, \
    LT_BSDF_FLAG_TRANSMISSION = 1, 2, 4, 8, 16
LT_LIGHT_SPOT, LT_LIGHT_AREA, LT_LIGHT_DIRECTIONAL, LT_LIGHT_POINT = 1, 2, 3, 4
LT_LIGHT_FINITE, LT_LIGHT_DELTA = 1 << 4, 1 << 5
LT_PTR_DEVICE = 1
LT_NEE_REFERENCE = 0
LT_NEE_CORRECTED = 1
LT_INTEGRATOR_PATH, LT_INTEGRATOR_BDPT, LT_INTEGRATOR_VCM, LT_INTEGRATOR_SPPM = 0, 1, 2, 3
LT_PATH_MEGAKERNEL, LT_PATH_WAVEFRONT, LT_PATH_PERSISTENT = 0, 1, 2
LT_MIS_BALANCE, LT_MIS_POWER = 0, 1
LT_VCM_REFERENCE, LT_VCM_CORRECTED = 0, 1
LT_WF_SORT_NONE, LT_WF_SORT_BSDF = 1, 2
LT_COLLAPSE_SAH, LT_COLLAPSE_GREEDY = 1, 2
LT_TREE_PLOC, LT_TREE_LBVH = 1, 2
LT_XCD_BANDS_OFF, LT_XCD_BANDS_ON = 1, 2
LT_SCENE_CORNELL, LT_SCENE_TERRAIN, LT_SCENE_CAUSTICS, LT_SCENE_DIFF_WALL = 0, 1, 2, 3


class lt_vertex(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("normal", C.c_float * 3), ("uv0", C.c_float * 2)]


class lt_prim_mesh_info(C.Structure):
    _fields_ = [("index_offset", C.c_uint32), ("vertex_offset", C.c_uint32), ("material_index", C.c_uint32),
                ("pad", C.c_uint32), ("min_pos", C.c_float * 4), ("max_pos", C.c_float * 4)]


class lt_material(C.Structure):
    _fields_ = [("albedo", C.c_float * 3), ("ior", C.c_float), ("emissive_factor", C.c_float * 3),
                ("bsdf_type", C.c_uint32), ("bsdf_props", C.c_uint32), ("k", C.c_float * 3),
                ("texture_id", C.c_int32), ("roughness", C.c_float), ("diffuse_trans", C.c_float),
                ("spec_trans", C.c_float), ("metallic", C.c_float), ("specular_tint", C.c_float),
                ("sheen_tint", C.c_float), ("clearcoat", C.c_float), ("clearcoat_gloss", C.c_float),
                ("sheen", C.c_float), ("subsurface", C.c_float), ("flatness", C.c_float),
                ("anisotropy", C.c_float), ("thin", C.c_uint32)]


class lt_light(C.Structure):
    _fields_ = [("world_matrix", C.c_float * 16), ("pos", C.c_float * 3), ("prim_mesh_idx", C.c_uint32),
                ("to", C.c_float * 3), ("num_triangles", C.c_uint32), ("L", C.c_float * 3),
                ("light_flags", C.c_uint32), ("world_center", C.c_float * 3), ("world_radius", C.c_float)]


class lt_texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("data", C.POINTER(C.c_float))]


class lt_scene_desc(C.Structure):
    _fields_ = [("vertices", C.POINTER(lt_vertex)), ("num_vertices", C.c_uint32),
                ("indices", C.POINTER(C.c_uint32)), ("num_indices", C.c_uint32),
                ("prim_infos", C.POINTER(lt_prim_mesh_info)), ("num_prim_infos", C.c_uint32),
                ("materials", C.POINTER(lt_material)), ("num_materials", C.c_uint32),
                ("lights", C.POINTER(lt_light)), ("num_lights", C.c_uint32),
                ("textures", C.POINTER(lt_texture)), ("num_textures", C.c_uint32),
                ("roughness_texture_ids", C.POINTER(C.c_int32)), ("spot_cones", C.POINTER(C.c_float)),
                ("total_light_triangle_cnt", C.c_uint32), ("total_light_area", C.c_float),
                ("dir_light_idx", C.c_int32), ("scene_min", C.c_float * 3), ("scene_max", C.c_float * 3),
                ("world_matrices", C.POINTER(C.c_float)), ("index_counts", C.POINTER(C.c_uint32))]


class lt_camera(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("projection", C.c_float * 16), ("inv_view", C.c_float * 16),
                ("inv_projection", C.c_float * 16)]


class lt_build_opts(C.Structure):
    _fields_ = [("max_leaf_tris", C.c_uint32), ("collapse", C.c_uint32), ("binary_tree", C.c_uint32),
                ("reserved", C.c_uint32 * 5)]


class lt_build_stats(C.Structure):
    _fields_ = [("num_nodes", C.c_uint32), ("num_tris", C.c_uint32), ("max_depth", C.c_uint32),
                ("binary_depth", C.c_uint32), ("build_ms", C.c_float)]


class lt_path_config(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_depth", C.c_int32),
                ("sky_col", C.c_float * 3), ("direct_lighting", C.c_uint32), ("strategy", C.c_uint32),
                ("nee_pick", C.c_uint32), ("reserved", C.c_uint32 * 5)]


class lt_bdpt_config(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_depth", C.c_int32), ("mis", C.c_uint32),
                ("reserved", C.c_uint32 * 8)]


class lt_vcm_config(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_depth", C.c_int32), ("sky_col", C.c_float * 3),
                ("radius_factor", C.c_float), ("alpha", C.c_float), ("enable_vm", C.c_uint32),
                ("estimator", C.c_uint32), ("reserved", C.c_uint32 * 6)]


class lt_sppm_config(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_depth", C.c_int32), ("sky_col", C.c_float * 3),
                ("base_radius", C.c_float), ("reserved", C.c_uint32 * 8)]


class lt_adjoint_config(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_depth", C.c_int32), ("spp", C.c_uint32),
                ("variants", C.c_uint32), ("albedo_texture", C.c_int32), ("roughness_texture", C.c_int32),
                ("sky_col", C.c_float * 3), ("reserved", C.c_uint32 * 8)]


class lt_frame_params(C.Structure):
    _fields_ = [("camera", lt_camera), ("frame_num", C.c_uint32), ("seed", C.c_uint32), ("spp", C.c_uint32),
                ("reserved", C.c_uint32 * 5)]


class lt_render_stats(C.Structure):
    _fields_ = [("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("nodes_visited", C.c_uint64),
                ("tris_tested", C.c_uint64), ("kernel_ms", C.c_float), ("frames", C.c_uint32),
                ("photons_tested", C.c_uint64), ("photons_gathered", C.c_uint64), ("gather_bytes", C.c_uint64),
                ("shadow_nodes_visited", C.c_uint64), ("shadow_tris_tested", C.c_uint64),
                ("num_phases", C.c_uint32), ("phase_ms", C.c_float * 8)]


class lt_procedural_params(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("terrain_tris", C.c_uint32), ("num_instances", C.c_uint32),
                ("num_point_lights", C.c_uint32), ("sphere_tris", C.c_uint32), ("wall_tris", C.c_uint32),
                ("tex_res", C.c_uint32), ("reserved", C.c_uint32 * 9)]


class lt_tuning(C.Structure):
    _fields_ = [("path_blocks_per_cu", C.c_uint32), ("trace_blocks_per_cu", C.c_uint32),
                ("refill_min", C.c_uint32), ("bdpt_blocks_per_cu", C.c_uint32), ("regen_min", C.c_uint32),
                ("bdpt_chunk_tiles", C.c_uint32), ("wf_sort", C.c_uint32), ("xcd_bands", C.c_uint32),
                ("reserved", C.c_uint32 * 8)]


class lt_scene_integrator_cfg(C.Structure):
    _fields_ = [("kind", C.c_int32), ("path_length", C.c_int32), ("sky_col", C.c_float * 3),
                ("enable_vm", C.c_int32), ("radius_factor", C.c_float), ("base_radius", C.c_float)]


# numpy views of the ray / hit records
RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("prim", "<u4"), ("inst", "<u4"), ("u", "<f4"), ("v", "<f4")])
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 20

P = C.c_void_p
_PROTOS = {
    "lt_context_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "lt_context_destroy": (C.c_int, [P]),
    "lt_last_error": (C.c_char_p, [P]),
    "lt_device_info": (C.c_int, [P, C.POINTER(C.c_int), C.c_char_p, C.c_size_t]),
    "lt_scene_upload": (C.c_int, [P, C.POINTER(lt_scene_desc)]),
    "lt_scene_validate": (C.c_int, [C.POINTER(lt_scene_desc), C.c_char_p, C.c_size_t]),
    "lt_accel_build": (C.c_int, [P, C.POINTER(lt_build_opts), C.POINTER(lt_build_stats)]),
    "lt_accel_export": (C.c_int, [P, P, C.c_size_t, P, C.c_size_t]),
    "lt_trace_closest": (C.c_int, [P, P, C.c_size_t, P, C.c_uint32]),
    "lt_trace_closest_noalloc": (C.c_int, [P, P, C.c_size_t, P]),
    "lt_trace_any": (C.c_int, [P, P, C.c_size_t, P, C.c_uint32]),
    "lt_trace_count": (C.c_int, [P, P, C.c_size_t, C.c_int, C.c_uint32, C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_uint64)]),
    "lt_last_kernel_ms": (C.c_int, [P, C.POINTER(C.c_float)]),
    "lt_set_tuning": (C.c_int, [P, C.POINTER(lt_tuning)]),
    "lt_get_tuning": (C.c_int, [P, C.POINTER(lt_tuning)]),
    "lt_device_alloc": (C.c_int, [P, C.c_size_t, C.POINTER(P)]),
    "lt_device_free": (C.c_int, [P, P]),
    "lt_memcpy_h2d": (C.c_int, [P, P, P, C.c_size_t]),
    "lt_memcpy_d2h": (C.c_int, [P, P, P, C.c_size_t]),
    "lt_synchronize": (C.c_int, [P]),
    "lt_stream_copy": (C.c_int, [P, P, P, C.c_size_t, C.POINTER(C.c_float)]),
    "lt_integrator_create": (C.c_int, [P, C.c_int, P, C.c_size_t, C.POINTER(P)]),
    "lt_integrator_destroy": (C.c_int, [P]),
    "lt_render": (C.c_int, [P, C.POINTER(lt_frame_params)]),
    "lt_render_stats_get": (C.c_int, [P, C.POINTER(lt_render_stats)]),
    "lt_render_set_timing": (C.c_int, [P, C.c_int]),
    "lt_render_set_counting": (C.c_int, [P, C.c_int]),
    "lt_render_set_graph": (C.c_int, [P, C.c_int]),
    "lt_accum_reset": (C.c_int, [P]),
    "lt_readback_rgba32f": (C.c_int, [P, P, C.c_size_t]),
    "lt_adjoint_create": (C.c_int, [P, C.POINTER(lt_adjoint_config), C.POINTER(P)]),
    "lt_adjoint_destroy": (C.c_int, [P]),
    "lt_adjoint_texture_size": (C.c_int, [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                          C.POINTER(C.c_uint32)]),
    "lt_adjoint_cache_floats": (C.c_int, [P, C.POINTER(C.c_size_t)]),
    "lt_adjoint_forward": (C.c_int, [P, C.POINTER(lt_camera), C.c_uint32, P, P, P, P, P]),
    "lt_adjoint_backward": (C.c_int, [P, C.POINTER(lt_camera), C.c_uint32, P, P, P, P, P, P, P]),
    "lt_adjoint_set_counting": (C.c_int, [P, C.c_int32]),
    "lt_adjoint_counts": (C.c_int, [P, C.POINTER(C.c_uint64)]),
    "lt_image_sq_error": (C.c_int, [P, P, P, C.c_size_t, C.c_uint32, C.POINTER(C.c_double)]),
    "lt_scene_procedural": (C.c_int, [C.c_int, C.POINTER(lt_procedural_params), C.POINTER(P)]),
    "lt_scene_load_json": (C.c_int, [C.c_char_p, C.POINTER(P), C.c_char_p, C.c_size_t]),
    "lt_scene_free": (C.c_int, [P]),
    "lt_scene_view": (C.c_int, [P, C.POINTER(lt_scene_desc)]),
    "lt_scene_camera": (C.c_int, [P, C.c_uint32, C.c_uint32, C.POINTER(lt_camera)]),
    "lt_scene_integrator": (C.c_int, [P, C.POINTER(lt_scene_integrator_cfg)]),
    "lt_scene_set_camera": (C.c_int, [P, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "lt_exr_save": (C.c_int, [C.c_char_p, P, C.c_uint32, C.c_uint32, C.c_int32, C.c_char_p, C.c_size_t]),
    "lt_exr_load": (C.c_int, [C.c_char_p, P, C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_char_p,
                              C.c_size_t]),
    "lt_image_load_rgba8": (C.c_int, [C.c_char_p, P, C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                      C.c_char_p, C.c_size_t]),
}
EXPORTED = tuple(_PROTOS)

_lib = None


def lib() -> C.CDLL:
    """The loaded liblt.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m lumen_amd.build` "
                              "(the HIP core has no CPU fallback)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _PROTOS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class LtError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def check(status: int, ctx=None, what: str = "") -> None:
    if status != LT_OK:
        msg = what
        if ctx is not None:
            msg = f"{what}: {lib().lt_last_error(ctx).decode()}"
        raise LtError(status, msg)
