"""Hand-built lt_scene_desc scenes for closed-form tests (host numpy).

Each scene is assembled directly in the reference wire layouts (Vertex,
PrimMeshInfo, Material, Light of src/shaders/commons.h), independently of
the product's scene generators, so analytic expectations can be stated.
"""
import ctypes as C

import numpy as np

from lumen_amd import _abi

PI = np.float32(3.14159265359)


class DescScene:
    def __init__(self):
        self.meshes = []  # (positions (T,3,3), normals (T,3,3) or None, material)
        self.materials = []
        self.analytic = []  # dict(type, pos, to, L)
        self.instances = []  # (mesh index of the geometry, material, 4x4 object-to-world)

    def material(self, **kw) -> int:
        m = _abi.lt_material()
        m.texture_id = -1
        for k, v in kw.items():
            if isinstance(v, (tuple, list)):
                getattr(m, k)[:] = list(v)
            else:
                setattr(m, k, v)
        self.materials.append(m)
        return len(self.materials) - 1

    def mesh(self, tris, mat, normals=None):
        tris = np.asarray(tris, np.float32).reshape(-1, 3, 3)
        if normals is None:
            n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
            n /= np.linalg.norm(n, axis=1, keepdims=True)
            normals = np.repeat(n[:, None, :], 3, axis=1)
        self.meshes.append((tris, np.asarray(normals, np.float32).reshape(-1, 3, 3), mat))

    def instance(self, mesh, mat, world):
        """Another mesh that shares mesh `mesh`'s geometry, with its own
        material and object-to-world matrix (4 x 4, row-major here)."""
        self.instances.append((mesh, mat, np.asarray(world, np.float64)))

    def build(self):
        verts, idx, prims, lights = [], [], [], []
        n_light_tris = 0
        counts, worlds = [], []
        for mi, (tris, nrm, mat) in enumerate(self.meshes):
            pi = _abi.lt_prim_mesh_info()
            pi.index_offset = len(idx)
            pi.vertex_offset = len(verts)
            pi.material_index = mat
            pi.min_pos[:3] = list(tris.reshape(-1, 3).min(0))
            pi.max_pos[:3] = list(tris.reshape(-1, 3).max(0))
            for t in range(len(tris)):
                for k in range(3):
                    v = _abi.lt_vertex()
                    v.pos[:] = list(tris[t, k])
                    v.normal[:] = list(nrm[t, k])
                    idx.append(3 * t + k)
                    verts.append(v)
            prims.append(pi)
            counts.append(3 * len(tris))
            worlds.append(np.eye(4))
        for mesh, mat, W in self.instances:
            src = prims[mesh]
            pi = _abi.lt_prim_mesh_info()
            pi.index_offset, pi.vertex_offset, pi.material_index = src.index_offset, src.vertex_offset, mat
            pi.min_pos[:3], pi.max_pos[:3] = list(src.min_pos[:3]), list(src.max_pos[:3])
            prims.append(pi)
            counts.append(counts[mesh])
            worlds.append(W)
        for mi, pi in enumerate(prims):
            m = self.materials[pi.material_index]
            if any(m.emissive_factor[i] > 0 for i in range(3)):
                L = _abi.lt_light()
                L.world_matrix[:] = list(worlds[mi].T.reshape(-1).astype(np.float32))  # column-major
                L.prim_mesh_idx = mi
                L.num_triangles = counts[mi] // 3
                L.light_flags = 2 | 16
                L.L[:] = list(m.emissive_factor)
                lights.append(L)
                n_light_tris += counts[mi] // 3
        allp = np.concatenate([t.reshape(-1, 3) for t, _, _ in self.meshes])
        lo, hi = allp.min(0), allp.max(0)
        for a in self.analytic:
            L = _abi.lt_light()
            L.pos[:] = list(a["pos"])
            L.to[:] = list(a["to"])
            L.L[:] = list(a["L"])
            L.light_flags = a["flags"]
            L.world_radius = float(np.linalg.norm(hi - lo) / 2)
            L.world_center[:] = list((lo + hi) / 2)
            lights.append(L)
            n_light_tris += 1
        self._v = (_abi.lt_vertex * len(verts))(*verts)
        self._i = (C.c_uint32 * len(idx))(*idx)
        self._p = (_abi.lt_prim_mesh_info * len(prims))(*prims)
        self._m = (_abi.lt_material * len(self.materials))(*self.materials)
        self._l = (_abi.lt_light * max(1, len(lights)))(*lights)
        d = _abi.lt_scene_desc()
        d.vertices, d.num_vertices = self._v, len(verts)
        d.indices, d.num_indices = self._i, len(idx)
        d.prim_infos, d.num_prim_infos = self._p, len(prims)
        d.materials, d.num_materials = self._m, len(self.materials)
        d.lights, d.num_lights = self._l, len(lights)
        d.total_light_triangle_cnt = n_light_tris
        d.dir_light_idx = -1
        d.scene_min[:] = list(lo)
        d.scene_max[:] = list(hi)
        if self.instances:
            self._w = (C.c_float * (16 * len(prims)))(*np.concatenate(
                [W.T.reshape(-1) for W in worlds]).astype(np.float32))
            self._c = (C.c_uint32 * len(prims))(*counts)
            d.world_matrices = C.cast(self._w, C.POINTER(C.c_float))
            d.index_counts = C.cast(self._c, C.POINTER(C.c_uint32))
        self.desc = d
        return d


def quad(a, b, c, d):
    return [[a, b, c], [a, c, d]]


def box_faces(lo, hi):
    """The 6 faces of an axis-aligned box, inward normals, each a quad."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    P = {k: np.array(v, np.float32) for k, v in dict(
        A=(x0, y0, z0), B=(x1, y0, z0), C=(x1, y0, z1), D=(x0, y0, z1),
        E=(x0, y1, z0), F=(x1, y1, z0), G=(x1, y1, z1), H=(x0, y1, z1)).items()}
    return [quad(P["A"], P["D"], P["C"], P["B"]), quad(P["E"], P["F"], P["G"], P["H"]),
            quad(P["A"], P["B"], P["F"], P["E"]), quad(P["D"], P["H"], P["G"], P["C"]),
            quad(P["A"], P["E"], P["H"], P["D"]), quad(P["B"], P["C"], P["G"], P["F"])]


def camera(pos, look_dir, fov_deg=40.0, aspect=1.0):
    """A pinhole camera in the SceneUBO convention the kernels read
    (sample_camera, commons.glsl:30-33): inv_view maps view to world, the
    view looks down -z; inv_projection maps NDC (x, y, 1, 1) to a view-space
    point.  Built here from first principles (orthonormal frame + tan(fov/2)).
    The inverse projection also maps NDC (x, y, 0, 1) to the pixel's point on
    the z = -1 plane after the divide by w (the camera-area computation of
    bdpt.rgen:48-51), and view / projection are its exact inverses, so the
    light-tracing raster projection (bdpt_commons.glsl:519-523) lands on the
    pixel whose camera ray points the same way."""
    f = np.asarray(look_dir, np.float64)
    f /= np.linalg.norm(f)
    up = np.array([0, 1, 0], np.float64) if abs(f[1]) < 0.99 else np.array([0, 0, 1], np.float64)
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    inv_view = np.eye(4)
    inv_view[:3, 0], inv_view[:3, 1], inv_view[:3, 2], inv_view[:3, 3] = r, u, -f, pos
    t = np.tan(np.radians(fov_deg) / 2)
    alpha = -0.5  # any non-zero value: only the w row at z = 1 depends on it
    inv_proj = np.diag([t * aspect, t, 0.0, 1.0])
    inv_proj[2, 3] = -1.0  # (x, y, z, 1) -> (x tan*a, y tan, -1, alpha z + 1)
    inv_proj[3, 2] = alpha
    proj = np.zeros((4, 4))
    proj[0, 0], proj[1, 1] = 1.0 / (t * aspect), 1.0 / t
    proj[2, 2] = proj[2, 3] = 1.0 / alpha
    proj[3, 2] = -1.0
    view = np.linalg.inv(inv_view)
    cam = _abi.lt_camera()
    cam.view[:] = list(view.T.reshape(-1).astype(np.float32))  # column-major
    cam.projection[:] = list(proj.T.reshape(-1).astype(np.float32))
    cam.inv_view[:] = list(inv_view.T.reshape(-1).astype(np.float32))
    cam.inv_projection[:] = list(inv_proj.T.reshape(-1).astype(np.float32))
    return cam
