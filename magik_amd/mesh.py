"""Triangle meshes for scenes with SCENE_MESHES: procedural meshes and the
conversion to the C-ABI structs of mgk_set_meshes (include/magik_amd.h).

A mesh is (vertices float32 (n, 3), indices uint32 (m, 3)) in object space,
SI metres; an instance is (mesh index, matter, 3 x 4 object -> world)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi


def box_mesh(lo, hi):
    """The 12 triangles of an axis-aligned box."""
    lo, hi = np.asarray(lo, np.float32), np.asarray(hi, np.float32)
    v = np.array([[(hi if (k >> a) & 1 else lo)[a] for a in range(3)] for k in range(8)], np.float32)
    faces = []
    for axis in range(3):
        for side in (0, 1):
            quad = [k for k in range(8) if ((k >> axis) & 1) == side]
            a, b, c, d = quad            # bits of the two other axes: 00, 01, 10, 11 in order
            faces += [(a, b, d), (a, d, c)]
    return v, np.array(faces, np.uint32)


def icosphere(subdivisions: int = 3, radius: float = 1.0):
    """A unit icosahedron subdivided `subdivisions` times, projected on the
    sphere: 20 * 4^subdivisions triangles."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = np.array(v, np.float64)
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    faces = np.array(f, np.int64)
    for _ in range(subdivisions):
        # one midpoint per edge, shared between the two faces of the edge
        e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
        key = np.sort(e, axis=1)
        uniq, inv = np.unique(key, axis=0, return_inverse=True)
        mid = verts[uniq[:, 0]] + verts[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(verts)
        verts = np.concatenate([verts, mid])
        n = len(faces)
        m01, m12, m20 = (base + inv.reshape(3, n)).astype(np.int64)
        a, b, c = faces[:, 0], faces[:, 1], faces[:, 2]
        faces = np.concatenate([np.stack([a, m01, m20], 1), np.stack([b, m12, m01], 1), np.stack([c, m20, m12], 1),
                                np.stack([m01, m12, m20], 1)])
    return (verts * radius).astype(np.float32), faces.astype(np.uint32)


def grid_plane(size: float, n: int, y: float = 0.0):
    """A square of side `size` in the plane y = const, centred on the y axis,
    split into n x n quads (2 n^2 triangles)."""
    g = np.linspace(-size / 2, size / 2, n + 1)
    xx, zz = np.meshgrid(g, g, indexing="ij")
    v = np.stack([xx.ravel(), np.full(xx.size, y), zz.ravel()], 1).astype(np.float32)
    i = np.arange(n)[:, None] * (n + 1) + np.arange(n)[None, :]
    a, b, c, d = i, i + 1, i + n + 1, i + n + 2
    f = np.concatenate([np.stack([a, b, d], -1).reshape(-1, 3), np.stack([a, d, c], -1).reshape(-1, 3)])
    return v, f.astype(np.uint32)


def affine(scale=1.0, translate=(0.0, 0.0, 0.0), rotate_y_deg: float = 0.0):
    """Row-major 3 x 4 object -> world map: rotation about y, then uniform or
    per-axis scale, then translation."""
    s = np.broadcast_to(np.asarray(scale, np.float64), (3,))
    th = np.radians(rotate_y_deg)
    R = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
    M = np.zeros((3, 4))
    M[:, :3] = np.diag(s) @ R
    M[:, 3] = translate
    return M


def to_abi(meshes, instances):
    """(Mesh array, Instance array, arrays to keep alive) for mgk_set_meshes."""
    keep = []
    m_arr = (abi.Mesh * max(1, len(meshes)))()
    for k, (v, f) in enumerate(meshes):
        v = np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3)
        f = np.ascontiguousarray(f, dtype=np.uint32).reshape(-1, 3)
        keep += [v, f]
        m_arr[k].vertices = v.ctypes.data_as(C.POINTER(C.c_float))
        m_arr[k].n_vertex = len(v)
        m_arr[k].n_triangle = len(f)
        m_arr[k].indices = f.ctypes.data_as(C.POINTER(C.c_uint32))
    i_arr = (abi.Instance * max(1, len(instances)))()
    for k, (mesh, matter, M) in enumerate(instances):
        i_arr[k].mesh = mesh
        i_arr[k].matter = matter
        for j, x in enumerate(np.asarray(M, np.float64).reshape(12)):
            i_arr[k].object_to_world[j] = x
    return m_arr, i_arr, keep


def hierarchy_info(meshes, instances):
    """(result code, MeshHierarchyInfo) of the hierarchy mgk_set_meshes would
    build, built and checked on the host (mgk_mesh_hierarchy_info)."""
    m, inst, _keep = to_abi(meshes, instances)
    info = abi.MeshHierarchyInfo()
    r = abi.load().mgk_mesh_hierarchy_info(m, len(meshes), inst, len(instances), C.byref(info))
    return r, info
