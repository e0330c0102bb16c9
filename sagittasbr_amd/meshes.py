"""Procedural PEC test bodies (vertices float32, triangle indices uint32).

Every generator returns ``(verts, faces)`` with ``verts`` of shape (V, 3)
float32 and ``faces`` of shape (F, 3) uint32, wound counter-clockwise seen
from the side the surface normal ``cross(v1-v0, v2-v0)`` points to. The
reference never flips normals (src/cuda/rtKernels.cu:42 and
src/cuda/poKernels.cu:94), so winding decides which side of a surface
reflects with a non-zero PO term; every generator here states its winding.

Geometry is computed in float64 and rounded to float32 once, matching the
float positions the reference's loaders hand to the triangle soup
(include/model/modelLoader.hpp:10-15, src/app/rtRcsEntry.cu:118-148).
"""
from __future__ import annotations

import math

import numpy as np

__all__ = [
    "icosphere",
    "plate",
    "box",
    "trihedral",
    "open_duct",
    "trihedral_duct_scene",
    "merge",
    "bounding_radius",
    "write_obj",
    "read_obj",
]


def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array(
        [
            [-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
            [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
            [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1],
        ],
        dtype=np.float64,
    )
    f = np.array(
        [
            [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
            [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
            [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
            [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
        ],
        dtype=np.int64,
    )
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v, f


def icosphere(subdivisions: int, radius: float = 1.0):
    """Icosphere by recursive 4:1 midpoint subdivision of an icosahedron.

    After each level the new vertices are pushed onto the unit sphere in
    float64; the result is scaled by ``radius`` and rounded to float32.
    Faces: 20 * 4**subdivisions; vertices: 10 * 4**subdivisions + 2.
    Winding: outward normals.
    """
    v, f = _icosahedron()
    for _ in range(int(subdivisions)):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid], axis=0)
        nf = f.shape[0]
        a = inv[0:nf] + base  # midpoint of (f0,f1)
        b = inv[nf:2 * nf] + base  # (f1,f2)
        c = inv[2 * nf:3 * nf] + base  # (f2,f0)
        f = np.concatenate(
            [
                np.stack([f[:, 0], a, c], axis=1),
                np.stack([f[:, 1], b, a], axis=1),
                np.stack([f[:, 2], c, b], axis=1),
                np.stack([a, b, c], axis=1),
            ],
            axis=0,
        )
    return (v * radius).astype(np.float32), f.astype(np.uint32)


def _grid_quad(origin, du, dv, nu, nv):
    """Planar nu x nv grid spanned by du, dv; normal = du x dv."""
    origin = np.asarray(origin, np.float64)
    du = np.asarray(du, np.float64)
    dv = np.asarray(dv, np.float64)
    s = np.arange(nu + 1, dtype=np.float64) / nu
    t = np.arange(nv + 1, dtype=np.float64) / nv
    S, T = np.meshgrid(s, t, indexing="ij")
    v = origin + S[..., None] * du + T[..., None] * dv
    v = v.reshape(-1, 3)
    idx = np.arange((nu + 1) * (nv + 1)).reshape(nu + 1, nv + 1)
    a = idx[:-1, :-1].ravel()
    b = idx[1:, :-1].ravel()
    c = idx[1:, 1:].ravel()
    d = idx[:-1, 1:].ravel()
    f = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)], axis=0)
    return v, f


def plate(size_u: float, size_v: float, nu: int = 1, nv: int = 1, z: float = 0.0):
    """Flat rectangle in the plane z=const, centred on the z axis, normal +z."""
    v, f = _grid_quad([-size_u / 2, -size_v / 2, z], [size_u, 0, 0], [0, size_v, 0], nu, nv)
    return v.astype(np.float32), f.astype(np.uint32)


def box(half: float = 1.0):
    """Axis-aligned cube of half-width ``half`` (12 triangles), outward normals."""
    h = half
    quads = [
        ([-h, -h, h], [2 * h, 0, 0], [0, 2 * h, 0]),  # +z
        ([-h, h, -h], [2 * h, 0, 0], [0, -2 * h, 0]),  # -z
        ([h, -h, -h], [0, 2 * h, 0], [0, 0, 2 * h]),  # +x
        ([-h, -h, -h], [0, 0, 2 * h], [0, 2 * h, 0]),  # -x
        ([-h, h, -h], [0, 0, 2 * h], [2 * h, 0, 0]),  # +y
        ([-h, -h, -h], [2 * h, 0, 0], [0, 0, 2 * h]),  # -y
    ]
    return merge([_to32(_grid_quad(o, du, dv, 1, 1)) for o, du, dv in quads])


def _to32(vf):
    v, f = vf
    return v.astype(np.float32), f.astype(np.uint32)


def merge(parts):
    """Concatenate (verts, faces) meshes into one soup."""
    vs, fs, base = [], [], 0
    for v, f in parts:
        vs.append(np.asarray(v, np.float32))
        fs.append(np.asarray(f, np.int64) + base)
        base += len(v)
    return np.concatenate(vs).astype(np.float32), np.concatenate(fs).astype(np.uint32)


def trihedral(a: float = 0.5, n: int = 64, corner=(0.0, 0.0, 0.0)):
    """Trihedral corner reflector: three mutually orthogonal a x a squares
    meeting at ``corner`` and spanning the +x/+y/+z octant. Each square is an
    n x n grid split into 2 triangles per cell (3*2*n*n triangles). Normals
    face into the octant, i.e. towards a radar at theta, phi in [0, 90] deg.
    """
    c = np.asarray(corner, np.float64)
    parts = [
        _grid_quad(c, [a, 0, 0], [0, a, 0], n, n),  # z = 0 floor, normal +z
        _grid_quad(c, [0, a, 0], [0, 0, a], n, n),  # x = 0 wall, normal +x
        _grid_quad(c, [0, 0, a], [a, 0, 0], n, n),  # y = 0 wall, normal +y
    ]
    return merge([_to32(p) for p in parts])


def open_duct(wx: float = 0.2, wy: float = 0.1, length: float = 0.6, n: int = 32, origin=(0.0, 0.0, 0.0)):
    """Rectangular duct of cross-section wx x wy and depth ``length`` along
    +z, open at z = origin.z + length and closed at z = origin.z. Five walls
    (4 sides + bottom), each an n x n grid split into triangles
    (5*2*n*n triangles), normals pointing into the cavity.
    """
    o = np.asarray(origin, np.float64)
    L = length
    parts = [
        _grid_quad(o, [wx, 0, 0], [0, wy, 0], n, n),  # bottom z=0, normal +z
        _grid_quad(o, [0, wy, 0], [0, 0, L], n, n),  # x=0 wall, normal +x
        _grid_quad(o + [wx, 0, 0], [0, 0, L], [0, wy, 0], n, n),  # x=wx wall, normal -x
        _grid_quad(o, [0, 0, L], [wx, 0, 0], n, n),  # y=0 wall, normal +y
        _grid_quad(o + [0, wy, 0], [wx, 0, 0], [0, 0, L], n, n),  # y=wy wall, normal -y
    ]
    return merge([_to32(p) for p in parts])


def trihedral_duct_scene():
    """The multi-bounce sweep scene: a 0.5 m trihedral (64x64 cells per
    face, 24,576 triangles) with its corner at the origin, plus a
    0.2 x 0.1 x 0.6 m duct (32x32 cells per wall, 10,240 triangles) beside it
    at x in [-0.45, -0.25], y in [0.2, 0.3], z in [-0.1, 0.5], open towards +z.
    34,816 triangles; every vertex lies within 0.8 m of the origin, so a
    1.6 m launch grid covers the body at every aspect.
    """
    tri = trihedral(0.5, 64)
    duct = open_duct(0.2, 0.1, 0.6, 32, origin=(-0.45, 0.2, -0.1))
    return merge([tri, duct])


def bounding_radius(verts) -> float:
    """Largest distance of a vertex from the origin (the launch grid is
    centred on the origin, src/app/sweep.cu:364)."""
    v = np.asarray(verts, np.float64)
    return float(np.sqrt((v * v).sum(1)).max()) if len(v) else 0.0


def write_obj(path, verts, faces):
    """Write a Wavefront OBJ with 9 significant digits, which round-trips
    float32 values exactly through a correctly rounded float parser."""
    v = np.asarray(verts, np.float32)
    f = np.asarray(faces, np.int64) + 1
    with open(path, "w") as fh:
        fh.write("# sagittasbr_amd procedural mesh\n")
        fh.writelines(f"v {a:.9g} {b:.9g} {c:.9g}\n" for a, b, c in v.tolist())
        fh.writelines(f"f {a} {b} {c}\n" for a, b, c in f.tolist())


def read_obj(path):
    """Minimal OBJ reader (v / f records, fan-triangulated polygons,
    1-based or negative indices)."""
    vs, fs = [], []
    with open(path) as fh:
        for line in fh:
            p = line.split()
            if not p:
                continue
            if p[0] == "v":
                vs.append([float(x) for x in p[1:4]])
            elif p[0] == "f":
                idx = []
                for tok in p[1:]:
                    i = int(tok.split("/")[0])
                    idx.append(i - 1 if i > 0 else len(vs) + i)
                for k in range(1, len(idx) - 1):
                    fs.append([idx[0], idx[k], idx[k + 1]])
    return np.asarray(vs, np.float32), np.asarray(fs, np.uint32).reshape(-1, 3)
