"""The C ABI (include/lt.h): the library loads and exports every declared
entry point; host-side scene and camera functions; error paths.  CPU only (no
compute call needs a GPU here)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import lumen_amd as la
from lumen_amd import _abi

HEADER = Path(__file__).resolve().parents[1] / "include" / "lt.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:lt_status|const char\*)\s+(lt_[a-z0-9_]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(_abi.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python mirror binds exactly the declared set
    assert sorted(_abi.EXPORTED) == names


def test_struct_layouts_match_reference_wire_format():
    assert C.sizeof(_abi.lt_vertex) == 32           # commons.h:194-198
    assert C.sizeof(_abi.lt_prim_mesh_info) == 48   # commons.h:333-340
    assert C.sizeof(_abi.lt_material) == 104        # commons.h:212-234
    assert C.sizeof(_abi.lt_light) == 128           # commons.h:200-210
    assert _abi.RAY_DTYPE.itemsize == 32 and _abi.HIT_DTYPE.itemsize == 20


def test_config_structs_match_the_header(tmp_path):
    """The ctypes mirrors of the config and parameter structs have the sizes
    and field offsets the C compiler gives the declarations of include/lt.h."""
    import shutil
    import subprocess

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = ["lt_path_config", "lt_bdpt_config", "lt_vcm_config", "lt_sppm_config", "lt_frame_params",
               "lt_render_stats", "lt_scene_integrator_cfg", "lt_adjoint_config", "lt_camera", "lt_scene_desc",
               "lt_tuning"]
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for name in structs:
        lines.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for fname, _ in getattr(_abi, name)._fields_:
            lines.append(f'printf("{name}.{fname} %zu\\n", offsetof({name}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "sizes.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "sizes"
    subprocess.run([cc, str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.splitlines())
    for name in structs:
        st = getattr(_abi, name)
        assert int(out[name]) == C.sizeof(st), name
        for fname, _ in st._fields_:
            assert int(out[f"{name}.{fname}"]) == getattr(st, fname).offset, f"{name}.{fname}"


def test_context_without_device_or_with_device():
    h = C.c_void_p()
    st = la.lib().lt_context_create(0, C.byref(h))
    assert st in (0, 7)  # LT_OK on a GPU box, LT_ERR_NO_DEVICE here
    if st == 0:
        la.lib().lt_context_destroy(h)
    assert la.lib().lt_context_destroy(None) == 0  # idempotent on NULL
    assert la.lib().lt_context_create(0, None) == 1


def test_cornell_scene_matches_its_definition():
    sc = la.Scene.procedural("cornell")
    d = sc.desc
    assert sc.num_triangles == 36
    assert d.num_lights == 1 and d.total_light_triangle_cnt == 2
    assert d.total_light_area == pytest.approx(0.25, rel=1e-6)
    L = d.lights[0]
    assert L.light_flags == (2 | 16) and tuple(L.L) == (17.0, 12.0, 4.0)
    mats = [d.materials[i] for i in range(d.num_materials)]
    albedos = sorted(tuple(round(x, 4) for x in m.albedo) for m in mats)
    assert (0.65, 0.05, 0.05) in albedos and (0.12, 0.45, 0.15) in albedos and (0.73, 0.73, 0.73) in albedos


def test_camera_closed_form_axis_aligned():
    """dir = +z: lookAtLH is a pure translation, the Euler round trip is the
    identity, inv_view = translate(pos) (Camera.h:28-38, 67-85) and the
    projection is Camera.h:109-116 (starting from the identity)."""
    sc = la.Scene.procedural("cornell")
    cam = sc.camera(256, 256)
    iv = np.array(cam.inv_view).reshape(4, 4)  # column-major rows = columns
    assert np.allclose(iv[:3, :3], np.eye(3)) and np.allclose(iv[3, :3], [0, 0, 3.9])
    P = np.array(cam.projection).reshape(4, 4)
    t = np.tan(np.radians(22.5))
    assert P[0, 0] == pytest.approx(1 / t, rel=1e-6) and P[1, 1] == pytest.approx(-1 / t, rel=1e-6)
    assert P[2, 3] == -1 and P[3, 3] == 1
    assert P[2, 2] == pytest.approx(1000 / (0.01 - 1000), rel=1e-6)
    # the centre ray looks down -z
    ip = np.array(cam.inv_projection).reshape(4, 4).T
    tgt = ip @ np.array([0, 0, 1, 1.0])
    dv = tgt[:3] / np.linalg.norm(tgt[:3])
    assert np.allclose(dv, [0, 0, -1], atol=1e-6)


def test_camera_round_trip_keeps_view_orthonormal():
    sc = la.Scene.procedural("caustics")
    for d in ([1, -2, 2], [-1.42, -2.16, 2], [0.3, 0.1, -1]):
        sc.set_camera(45, [1, 2, 3], d)
        iv = np.array(sc.camera(64, 48).inv_view).reshape(4, 4)[:3, :3]
        assert np.allclose(iv @ iv.T, np.eye(3), atol=1e-5)
        assert np.linalg.det(iv) == pytest.approx(1.0, abs=1e-5)


@pytest.mark.gpu
def test_scene_upload_rejects_out_of_range_arrays():
    # host-side validation runs before any device copy or kernel
    s = la.Scene.procedural("cornell")
    h = C.c_void_p()
    assert la.lib().lt_context_create(0, C.byref(h)) == 0
    try:
        d = _abi.lt_scene_desc.from_buffer_copy(s.desc)
        d.num_indices = 35
        assert la.lib().lt_scene_upload(h, C.byref(d)) == 1
        d = _abi.lt_scene_desc.from_buffer_copy(s.desc)
        d.num_materials = 1
        assert la.lib().lt_scene_upload(h, C.byref(d)) == 1
        assert b"out of range" in la.lib().lt_last_error(h)
    finally:
        la.lib().lt_context_destroy(h)


def test_terrain_workload_size():
    sc = la.Scene.procedural("terrain", seed=1, terrain_tris=20000, num_instances=10, num_point_lights=4)
    d = sc.desc
    assert d.num_lights == 1 + 4
    assert d.total_light_triangle_cnt == 2 + 4  # each analytic light counts one (LumenScene.cpp:106)
    assert sc.num_triangles == 2 * 100 * 100 + 10 * 12 + 2
    # the boxes are instances of one 12-triangle cube: its geometry is stored once
    assert d.num_indices // 3 == 2 * 100 * 100 + 12 + 2
    assert d.num_prim_infos == 16 + 10 + 1 and bool(d.world_matrices) and bool(d.index_counts)


def _validate(desc):
    err = C.create_string_buffer(256)
    st = la.lib().lt_scene_validate(C.byref(desc), err, 256)
    return st, err.value.decode()


@pytest.mark.parametrize("kind,kw", [("cornell", {}), ("caustics", dict(sphere_tris=2000)), ("diff_wall", {}),
                                     ("terrain", dict(terrain_tris=2000, num_instances=20, num_point_lights=4))])
def test_procedural_scenes_pass_upload_validation(kind, kw):
    """lt_scene_validate runs the host-side checks of lt_scene_upload (index
    ranges, references, instance transforms) without a device."""
    sc = la.Scene.procedural(kind, seed=1, **kw)
    assert _validate(sc.desc) == (0, "")


def test_upload_validation_refusals():
    from scenes import DescScene, quad
    s = DescScene()
    m = s.material(bsdf_type=1, bsdf_props=9, albedo=(0.5, 0.5, 0.5))
    e = s.material(bsdf_type=1, bsdf_props=9, albedo=(1, 1, 1), emissive_factor=(2, 2, 2))
    s.mesh(quad([-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]), m)
    W = np.eye(4)
    W[:3, 3] = (0, 0, 2)
    s.instance(0, e, W)
    s.analytic.append(dict(pos=(0, 3, 0), to=(0, 0, 0), L=(5, 5, 5), flags=4 | 16 | 32))  # a point light
    d = s.build()
    assert _validate(d) == (0, "")  # a point light's unset world matrix is not read
    s._w[16 * 1 + 0] = 0.0  # instance 1: a singular matrix (x column zero)
    st, msg = _validate(d)
    assert st == 1 and "world_matrices[1]" in msg
    s._w[16 * 1 + 0] = 1.0
    s._c[1] = 9  # more indices than the geometry has
    st, msg = _validate(d)
    assert st == 1 and "prim_infos[1]" in msg
    s._c[1] = 6
    d.lights[0].num_triangles = 3  # the instance's light claims 3 triangles of 2
    st, msg = _validate(d)
    assert st == 1 and "triangle count" in msg
    d.lights[0].num_triangles = 2
    assert _validate(d) == (0, "")
    for bad in (float("nan"), float("inf")):  # a poisoned vertex position
        old = d.vertices[2].pos[1]
        d.vertices[2].pos[1] = bad
        st, msg = _validate(d)
        assert st == 1 and "vertices[2].pos" in msg
        d.vertices[2].pos[1] = old
    assert _validate(d) == (0, "")
