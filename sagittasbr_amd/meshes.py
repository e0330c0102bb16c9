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
    "airframe",
    "ellipsoid",
    "cylinder",
    "random_bodies",
    "sh_perturbed_sphere",
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


# ------------------------------------------------------------------ airframe
def _naca00(x, t=0.12):
    """NACA 4-digit symmetric half-thickness at chord fraction x (closed TE)."""
    return 5.0 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3 - 0.1015 * x ** 4)


def _lifting_surface(root_le, span_dir, span, c_root, c_tip, sweep_deg, n_chord, n_span, thick_axis, t=0.12):
    """Trapezoidal lifting surface (wing / tail) with a NACA-00xx section,
    as a closed loop of n_chord points around the section times n_span + 1
    stations; outward normals. Triangles: 2 * n_chord * n_span."""
    root_le = np.asarray(root_le, np.float64)
    s_dir = np.asarray(span_dir, np.float64)
    t_dir = np.asarray(thick_axis, np.float64)
    c_dir = np.cross(s_dir, t_dir)  # chordwise (leading -> trailing edge)
    if c_dir @ np.array([1.0, 0.0, 0.0]) < 0:
        c_dir = -c_dir
    # cosine spacing around the loop: upper surface LE->TE, lower TE->LE
    half = n_chord // 2
    beta = np.linspace(0.0, np.pi, half + 1)
    xc = 0.5 * (1 - np.cos(beta))
    loop_x = np.concatenate([xc[:-1], xc[::-1][:-1]])  # n_chord points
    loop_s = np.concatenate([np.ones(half), -np.ones(half)])
    eta = np.arange(n_span + 1) / n_span
    chord = c_root + (c_tip - c_root) * eta
    le_off = span * eta * math.tan(math.radians(sweep_deg))
    X = loop_x[None, :] * chord[:, None]
    Y = loop_s[None, :] * _naca00(loop_x, t)[None, :] * chord[:, None]
    P = (root_le[None, None, :] + (le_off[:, None, None] + X[..., None]) * c_dir
         + (span * eta)[:, None, None] * s_dir + Y[..., None] * t_dir)
    v = P.reshape(-1, 3)
    nc = n_chord
    i = np.arange(n_span)[:, None]
    j = np.arange(nc)[None, :]
    a = (i * nc + j).ravel()
    b = (i * nc + (j + 1) % nc).ravel()
    c = ((i + 1) * nc + (j + 1) % nc).ravel()
    d = ((i + 1) * nc + j).ravel()
    f = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    # orient outward: check against the section centroid line
    cen = v[f].mean(1)
    nrm = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    axis_pt = root_le + (span * eta.mean()) * s_dir
    rel = cen - axis_pt
    rel -= np.outer(rel @ s_dir, s_dir)
    rel -= np.outer(rel @ c_dir, c_dir)
    if np.mean(np.einsum("ij,ij->i", nrm, rel)) < 0:
        f = f[:, [0, 2, 1]]
    return v, f


def airframe(seed: int = 1, n_fuse=(2048, 512), n_wing=(1024, 256), n_fin=(256, 64)):
    """Procedural airframe, about 3.2M triangles at the default resolution.

    Fuselage: 20 m long along +x, lofted from elliptical sections (semi-axes
    r(s) and 1.1 r(s)) with r rising to 1.2 m (ogive nose, cylindrical body,
    tapered tail), n_fuse = (stations, points per section), 2 triangles per
    cell (2,097,152). Wings: two swept trapezoidal panels of 8 m semi-span
    (16 m span) with a NACA-0012 section, n_wing = (points around the
    section, span stations) each (1,048,576). Tail: one vertical and two
    horizontal fins of n_fin cells each. The seed perturbs sweep, chords and
    dihedral within a few percent. Outward normals, centred on the origin.
    """
    rng = np.random.default_rng(seed)
    L, R = 20.0, 1.2
    ns, nr = n_fuse
    s = (np.arange(ns + 1) + 0.5) / (ns + 1)
    nose, tail = 0.18, 0.72
    r = np.where(s < nose, R * np.sqrt(np.clip(1 - ((nose - s) / nose) ** 2, 0, 1)),
                 np.where(s > tail, R * (1 - 0.8 * np.clip((s - tail) / (1 - tail), 0, 1) ** 1.5), R))
    r = np.maximum(r, 0.02)
    th = 2 * np.pi * np.arange(nr) / nr
    X = (s * L - L / 2)[:, None] * np.ones_like(th)[None, :]
    Y = r[:, None] * np.cos(th)[None, :]
    Z = 1.1 * r[:, None] * np.sin(th)[None, :]
    fv = np.stack([X, Y, Z], -1).reshape(-1, 3)
    i = np.arange(ns)[:, None]
    j = np.arange(nr)[None, :]
    a = (i * nr + j).ravel()
    b = ((i + 1) * nr + j).ravel()
    c = ((i + 1) * nr + (j + 1) % nr).ravel()
    d = (i * nr + (j + 1) % nr).ravel()
    ff = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    # ring k has normal ~ (0, cos, 1.1 sin): (a,b,c) winds x then theta -> outward
    nrm = np.cross(fv[ff[:, 1]] - fv[ff[:, 0]], fv[ff[:, 2]] - fv[ff[:, 0]])
    rad = fv[ff].mean(1).copy()
    rad[:, 0] = 0
    if np.mean(np.einsum("ij,ij->i", nrm, rad)) < 0:
        ff = ff[:, [0, 2, 1]]
    parts = [(fv, ff)]
    sweep = 28.0 * (1 + 0.05 * rng.uniform(-1, 1))
    c_root = 3.6 * (1 + 0.05 * rng.uniform(-1, 1))
    c_tip = 1.3 * (1 + 0.05 * rng.uniform(-1, 1))
    dih = math.radians(4.0 * (1 + 0.1 * rng.uniform(-1, 1)))
    nc, nsp = n_wing
    for side in (1.0, -1.0):
        sd = np.array([0.0, side * math.cos(dih), math.sin(dih)])
        parts.append(_lifting_surface([-1.5, side * 1.05, -0.3], sd, 8.0, c_root, c_tip, sweep, nc, nsp,
                                      thick_axis=[0.0, -side * math.sin(dih), math.cos(dih)]))
    fc, fs = n_fin
    parts.append(_lifting_surface([6.3, 0.0, 0.9], [0.0, 0.0, 1.0], 3.2, 2.6, 1.1, 40.0, fc, fs,
                                  thick_axis=[0.0, 1.0, 0.0], t=0.10))
    for side in (1.0, -1.0):
        parts.append(_lifting_surface([6.6, side * 0.4, 0.2], [0.0, side, 0.0], 3.4, 2.0, 0.9, 32.0, fc, fs,
                                      thick_axis=[0.0, 0.0, 1.0], t=0.10))
    return merge([(p[0].astype(np.float32), p[1].astype(np.uint32)) for p in parts])


# ---------------------------------------------------------- random bodies
def _uv_sphere_counts(n_tris):
    """(rings, segments) of a UV-sphere-like tessellation with ~n_tris."""
    seg = max(8, int(round(math.sqrt(n_tris))))
    rings = max(4, int(round(n_tris / (2 * seg))))
    return rings, seg


def ellipsoid(a, b, c, n_tris):
    """Ellipsoid with semi-axes a, b, c; latitude/longitude grid with ~n_tris
    triangles (caps included); outward normals."""
    rings, seg = _uv_sphere_counts(n_tris)
    th = np.pi * (np.arange(1, rings) / rings)
    ph = 2 * np.pi * np.arange(seg) / seg
    T, P = np.meshgrid(th, ph, indexing="ij")
    v = np.stack([a * np.sin(T) * np.cos(P), b * np.sin(T) * np.sin(P), c * np.cos(T)], -1).reshape(-1, 3)
    v = np.concatenate([v, [[0, 0, c], [0, 0, -c]]])
    top, bot = len(v) - 2, len(v) - 1
    i = np.arange(rings - 2)[:, None]
    j = np.arange(seg)[None, :]
    a0 = (i * seg + j).ravel()
    b0 = (i * seg + (j + 1) % seg).ravel()
    c0 = ((i + 1) * seg + (j + 1) % seg).ravel()
    d0 = ((i + 1) * seg + j).ravel()
    f = [np.stack([a0, d0, c0], 1), np.stack([a0, c0, b0], 1)]
    jj = np.arange(seg)
    f.append(np.stack([np.full(seg, top), jj, (jj + 1) % seg], 1))
    last = (rings - 2) * seg
    f.append(np.stack([np.full(seg, bot), last + (jj + 1) % seg, last + jj], 1))
    return v.astype(np.float32), np.concatenate(f).astype(np.uint32)


def cylinder(radius, height, n_tris, cavity=None):
    """Closed finite cylinder along z (centred), ~n_tris triangles, outward
    normals. ``cavity=(w, h, depth)`` cuts a rectangular pocket into the +z
    cap: the cap becomes a frame around a w x h opening whose walls and floor
    face into the pocket."""
    seg = max(12, int(round(math.sqrt(n_tris / 2))))
    nz = max(2, int(round(n_tris / (2 * seg) * 0.8)))
    ph = 2 * np.pi * np.arange(seg) / seg
    z = height * (np.arange(nz + 1) / nz - 0.5)
    P, Zg = np.meshgrid(ph, z, indexing="ij")
    side = np.stack([radius * np.cos(P), radius * np.sin(P), Zg], -1).reshape(-1, 3)
    i = np.arange(seg)[:, None]
    k = np.arange(nz)[None, :]
    a0 = (i * (nz + 1) + k).ravel()
    b0 = (((i + 1) % seg) * (nz + 1) + k).ravel()
    c0 = (((i + 1) % seg) * (nz + 1) + k + 1).ravel()
    d0 = (i * (nz + 1) + k + 1).ravel()
    parts = [(side, np.concatenate([np.stack([a0, b0, c0], 1), np.stack([a0, c0, d0], 1)]))]
    # caps as triangle fans from the centre, refined in rings
    nring = max(1, seg // 8)
    for zc, up in ((height / 2, True), (-height / 2, False)):
        rr = radius * np.arange(1, nring + 1) / nring
        R_, P_ = np.meshgrid(rr, ph, indexing="ij")
        cv = np.stack([R_ * np.cos(P_), R_ * np.sin(P_), np.full_like(R_, zc)], -1).reshape(-1, 3)
        cv = np.concatenate([[[0, 0, zc]], cv])
        jj = np.arange(seg)
        fs = [np.stack([np.zeros(seg, int), 1 + jj, 1 + (jj + 1) % seg], 1)]
        for q in range(nring - 1):
            a1 = 1 + q * seg + jj
            b1 = 1 + q * seg + (jj + 1) % seg
            c1 = 1 + (q + 1) * seg + (jj + 1) % seg
            d1 = 1 + (q + 1) * seg + jj
            fs += [np.stack([a1, d1, c1], 1), np.stack([a1, c1, b1], 1)]
        fc = np.concatenate(fs)
        if not up:
            fc = fc[:, [0, 2, 1]]
        parts.append((cv, fc))
    v, f = merge([(p[0].astype(np.float32), p[1].astype(np.uint32)) for p in parts])
    if cavity is not None:
        w, h, depth = cavity
        # pocket: 4 walls + floor below the +z cap, normals into the pocket
        z0 = height / 2
        pk = [
            _grid_quad([-w / 2, -h / 2, z0 - depth], [w, 0, 0], [0, h, 0], 4, 4),  # floor, +z
            _grid_quad([-w / 2, -h / 2, z0 - depth], [0, h, 0], [0, 0, depth], 4, 4),  # +x facing
            _grid_quad([w / 2, -h / 2, z0 - depth], [0, 0, depth], [0, h, 0], 4, 4),  # -x facing
            _grid_quad([-w / 2, -h / 2, z0 - depth], [0, 0, depth], [w, 0, 0], 4, 4),  # +y facing
            _grid_quad([-w / 2, h / 2, z0 - depth], [w, 0, 0], [0, 0, depth], 4, 4),  # -y facing
        ]
        v, f = merge([(v, f)] + [_to32(q) for q in pk])
    return v, f


def random_bodies(seed: int = 2, n: int = 64, tri_range=(5_000, 400_000)):
    """Seeded batch of PEC bodies: ellipsoids (semi-axes U[0.2, 1.5] m),
    finite cylinders and cylinders with a rectangular cavity, each with a
    log-uniform triangle budget in ``tri_range`` and its own frequency
    f ~ U[1, 10] GHz. Returns a list of dicts (verts, faces, freq, kind)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nt = int(round(math.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1])))))
        kind = ("ellipsoid", "cylinder", "cavity")[k % 3]
        if kind == "ellipsoid":
            a, b, c = rng.uniform(0.2, 1.5, 3)
            v, f = ellipsoid(a, b, c, nt)
        else:
            rad = rng.uniform(0.2, 0.8)
            hgt = rng.uniform(0.4, 2.0)
            cav = (rad * 0.8, rad * 0.5, hgt * 0.4) if kind == "cavity" else None
            v, f = cylinder(rad, hgt, nt, cav)
        out.append(dict(verts=v, faces=f, freq=float(rng.uniform(1e9, 10e9)), kind=kind))
    return out


# ------------------------------------------------ SH-perturbed icosphere
def sh_perturbed_sphere(subdivisions: int = 7, radius: float = 1.0, degree: int = 8, rms: float = 0.02,
                        seed: int = 3):
    """Icosphere whose vertices move radially by a seeded sum of real
    spherical harmonics of degree 1..``degree`` scaled to ``rms`` metres RMS
    displacement over the vertices. Returns (verts, faces, displacement)."""
    from scipy.special import sph_harm_y

    v, f = icosphere(subdivisions, 1.0)
    u = v.astype(np.float64)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    theta = np.arccos(np.clip(u[:, 2], -1, 1))
    phi = np.arctan2(u[:, 1], u[:, 0])
    rng = np.random.default_rng(seed)
    disp = np.zeros(len(u))
    for l in range(1, degree + 1):
        for m in range(-l, l + 1):
            y = sph_harm_y(l, abs(m), theta, phi)
            basis = np.sqrt(2) * (y.imag if m < 0 else y.real) if m != 0 else y.real
            disp += rng.normal() / (l + 1) * basis
    disp *= rms / np.sqrt(np.mean(disp ** 2))
    pts = u * (radius + disp)[:, None]
    return pts.astype(np.float32), f, disp
