"""ctypes access to the CPU oracle (oracle/liboracle.so) and to the
reference BVH builder compiled from the reference sources
(oracle/_ref/libref_bvh_f{32,64}.so). Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = {
    "f32": os.path.join(ROOT, "oracle", "_ref", "libref_bvh_f32.so"),
    "f64": os.path.join(ROOT, "oracle", "_ref", "libref_bvh_f64.so"),
}

_P = C.c_void_p
_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run `make -C oracle`)")
        L = C.CDLL(ORACLE_SO)
        L.orc_scene_create.restype = _P
        L.orc_scene_create.argtypes = [_fp, C.c_int, _up, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_scene_destroy.argtypes = [_P]
        L.orc_scene_info.argtypes = [_P, _ip]
        L.orc_scene_export.argtypes = [_P, _dp, _ip, _dp, _ip]
        L.orc_frame.argtypes = [C.c_int, _dp, _ip, C.c_int, _dp, _dp]
        L.orc_trace_angle.argtypes = [_P, _dp, _ip, C.c_int, C.c_int, _P, _P, _P, _P, _P, _P]
        L.orc_trace_angle.restype = C.c_int
        L.orc_sigma.argtypes = [C.c_int, C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_trace_one.argtypes = [_P, _dp, _dp, C.c_int, C.c_int, _ip, C.POINTER(C.c_double)]
        L.orc_trace_one.restype = C.c_int
        L.orc_closest_pair.argtypes = [_P, _dp, _dp, _ip, _dp]
        L.orc_hit_triangle.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, _dp]
        L.orc_hit_triangle.restype = C.c_int
        L.orc_sweep.argtypes = [_P, _dp, _ip, C.c_int, _ip, C.c_int, _dp, _lp]
        L.orc_sweep.restype = C.c_int
        L.orc_num_threads.restype = C.c_int
        L.orc_frozen_sum.argtypes = [_dp, _up, _dp, _ip, _ip, _ip, _dp]
        L.orc_frozen_sum.restype = C.c_int
        L.orc_frozen_fd.argtypes = [_dp, C.c_int, _up, C.c_int, C.c_int, _dp, _ip, _ip, _ip, _dp, C.c_int, _ip,
                                    C.c_double, C.c_int, _dp, _dp]
        L.orc_frozen_fd.restype = C.c_int
        _lib = L
    return _lib


@dataclass
class Params:
    """Mirror of the reference's simulationConfig (include/app/config.hpp:7-27)."""

    theta_start: float = 90.0
    theta_end: float = 90.0
    theta_samples: int = 1
    phi_start: float = 0.0
    phi_end: float = 180.0
    phi_samples: int = 181
    grid_size: float = 10.0
    n_x: int = 5000
    n_y: int = 5000
    max_bounces: int = 20
    reflection_const: float = 1.0
    freq: float = 10.0e9

    def pd(self):
        return np.array(
            [self.theta_start, self.theta_end, self.phi_start, self.phi_end, self.grid_size, self.freq,
             self.reflection_const], np.float64)

    def pi(self):
        return np.array([self.theta_samples, self.phi_samples, self.n_x, self.n_y, self.max_bounces], np.int32)

    @property
    def n_angles(self):
        return self.theta_samples * self.phi_samples

    @property
    def n_rays(self):
        return self.n_x * self.n_y


class Scene:
    """An oracle scene: triangle soup + BVH in FP32 (precision=0) or FP64 (1)."""

    def __init__(self, verts, faces, precision=1, scale=1.0, leaf_size=4, bin_count=16, sah=True):
        v = np.ascontiguousarray(verts, np.float32).reshape(-1)
        f = np.ascontiguousarray(faces, np.uint32).reshape(-1)
        self.precision = precision
        self.h = lib().orc_scene_create(v, len(v) // 3, f, len(f) // 3, float(scale), precision, leaf_size,
                                        bin_count, int(sah))
        if not self.h:
            raise ValueError("oracle rejected the mesh (index out of range)")
        info = np.zeros(4, np.int32)
        lib().orc_scene_info(self.h, info)
        self.n_tris, self.n_nodes, self.root, self.depth = (int(x) for x in info)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def export(self):
        tris = np.zeros((self.n_tris, 9), np.float64)
        perm = np.zeros(self.n_tris, np.int32)
        bounds = np.zeros((self.n_nodes, 6), np.float64)
        links = np.zeros((self.n_nodes, 4), np.int32)
        lib().orc_scene_export(self.h, tris, perm, bounds, links)
        return tris, perm, bounds, links

    def trace_angle(self, p: Params, angle: int, fused=False, ids=True, per_ray=True):
        n = p.n_rays
        out = {}
        idb = np.full(n * max(p.max_bounces, 1), -1, np.int32) if ids else None
        cnt = np.zeros(n, np.int32) if per_ray else None
        L = np.zeros(n, np.float64) if per_ray else None
        reim = np.zeros(2 * n, np.float64) if per_ray else None
        s2 = np.zeros(2, np.float64)
        c3 = np.zeros(3, np.int64)
        ptr = lambda a: a.ctypes.data if a is not None else None
        rc = lib().orc_trace_angle(self.h, p.pd(), p.pi(), angle, int(fused), ptr(idb), ptr(cnt), ptr(L), ptr(reim),
                                   ptr(s2), ptr(c3))
        if rc:
            raise ValueError("orc_trace_angle rejected its arguments")
        if ids:
            out["ids"] = idb.reshape(n, -1)[:, : p.max_bounces]
        if per_ray:
            out["count"] = cnt
            out["L"] = L
            out["reim"] = reim.reshape(n, 2)
        out["sum"] = complex(s2[0], s2[1])
        out["visits"], out["tests"], out["drops"] = (int(x) for x in c3)
        return out

    def sweep(self, p: Params, angles=None, fused=False):
        a = np.arange(p.n_angles, dtype=np.int32) if angles is None else np.ascontiguousarray(angles, np.int32)
        reim = np.zeros(2 * len(a), np.float64)
        c3 = np.zeros(3, np.int64)
        rc = lib().orc_sweep(self.h, p.pd(), p.pi(), int(fused), a, len(a), reim, c3)
        if rc:
            raise ValueError("orc_sweep rejected its arguments")
        return reim[0::2] + 1j * reim[1::2], c3

    def trace_one(self, o, d, max_bounces, fused=False):
        ids = np.full(max(max_bounces, 1), -1, np.int32)
        L = C.c_double(0.0)
        n = lib().orc_trace_one(self.h, np.asarray(o, np.float64), np.asarray(d, np.float64), max_bounces,
                                int(fused), ids, C.byref(L))
        return n, ids[:max_bounces], L.value

    def closest_pair(self, o, d):
        out = np.zeros(2, np.int32)
        t = np.zeros(2, np.float64)
        lib().orc_closest_pair(self.h, np.asarray(o, np.float64), np.asarray(d, np.float64), out, t)
        return (int(out[0]), t[0]), (int(out[1]), t[1])


def frame(p: Params, angle: int, precision=1):
    o12 = np.zeros(12, np.float64)
    o4 = np.zeros(4, np.float64)
    lib().orc_frame(precision, p.pd(), p.pi(), angle, o12, o4)
    return o12.reshape(4, 3), o4


def sigma(re, im, precision=1):
    s = C.c_double()
    d = C.c_double()
    lib().orc_sigma(precision, float(re), float(im), C.byref(s), C.byref(d))
    return s.value, d.value


def hit_triangle(v9, o, d, tmin=1e-3, tmax=1e20, precision=1):
    tuv = np.zeros(3, np.float64)
    h = lib().orc_hit_triangle(precision, np.asarray(v9, np.float64).reshape(-1), np.asarray(o, np.float64),
                               np.asarray(d, np.float64), tmin, tmax, tuv)
    return bool(h), tuv


def ref_available(prec="f64"):
    return os.path.exists(REF_SO[prec])


def ref_bvh(verts, faces, prec="f64", parallel=False, scale=1.0):
    """Run the reference's buildBvh on the triangle soup; returns
    (tris9 in BVH order, bounds (N,6), links (N,4), root)."""
    L = C.CDLL(REF_SO[prec])
    L.ref_bvh_build.restype = C.c_int
    L.ref_bvh_build.argtypes = [_fp, C.c_int, C.c_int, _dp, _dp, _ip, C.POINTER(C.c_int)]
    v = np.asarray(verts, np.float32)
    f = np.asarray(faces, np.int64)
    soup = (v[f].reshape(-1, 9).astype(np.float64) * scale)
    soup32 = soup.astype(np.float32) if scale == 1.0 else v[f].reshape(-1, 9).astype(np.float32)
    n = len(f)
    tris = np.zeros((n, 9), np.float64)
    bounds = np.zeros((max(2 * n - 1, 1), 6), np.float64)
    links = np.zeros((max(2 * n - 1, 1), 4), np.int32)
    root = C.c_int(0)
    nn = L.ref_bvh_build(np.ascontiguousarray(soup32.reshape(-1)), n, int(parallel), tris, bounds, links,
                         C.byref(root))
    return tris, bounds[:nn], links[:nn], root.value


def canonical_tree(bounds, links, root):
    """Pre-order (left before right) renumbering of a BVH so trees built
    with different node numberings compare equal."""
    order = []
    stack = [root]
    while stack:
        n = stack.pop()
        order.append(n)
        l, r, ft, tc = links[n]
        if tc == 0 and l >= 0:
            stack.append(r)
            stack.append(l)
    pos = {n: i for i, n in enumerate(order)}
    cb = bounds[order]
    cl = links[order].copy()
    for i, n in enumerate(order):
        l, r, ft, tc = links[n]
        if tc == 0 and l >= 0:
            cl[i, 0] = pos[l]
            cl[i, 1] = pos[r]
    return cb, cl


# ------------------------------------------------------------ frozen paths
def job_arrays(theta_deg, phi_deg, grid_size, freq, n_x, n_y, max_bounces, reflection_const=1.0):
    """(pj, ij) arrays of the frozen-path functions for explicit directions."""
    th = np.atleast_1d(np.asarray(theta_deg, np.float64))
    ph = np.atleast_1d(np.asarray(phi_deg, np.float64))
    fr = np.broadcast_to(np.asarray(freq, np.float64), th.shape)
    pj = np.stack([th, ph, np.full_like(th, grid_size), fr, np.full_like(th, reflection_const)], 1)
    ij = np.tile(np.array([n_x, n_y, max_bounces], np.int32), (len(th), 1))
    return np.ascontiguousarray(pj), np.ascontiguousarray(ij)


def direction_params(theta, phi, grid_size, freq, n_x, n_y, max_bounces, reflection_const=1.0):
    """A one-angle sweep whose frame is the frame of direction (theta, phi)."""
    return Params(theta_start=float(theta), theta_end=float(theta), theta_samples=1, phi_start=float(phi),
                  phi_end=float(phi), phi_samples=1, grid_size=grid_size, n_x=n_x, n_y=n_y,
                  max_bounces=max_bounces, reflection_const=reflection_const, freq=freq)


def frozen_sum(verts64, faces, pj, ij, ids, cnt):
    out = np.zeros(2, np.float64)
    rc = lib().orc_frozen_sum(np.ascontiguousarray(verts64, np.float64).reshape(-1),
                              np.ascontiguousarray(faces, np.uint32).reshape(-1), np.ascontiguousarray(pj, np.float64),
                              np.ascontiguousarray(ij, np.int32), np.ascontiguousarray(ids, np.int32).reshape(-1),
                              np.ascontiguousarray(cnt, np.int32).reshape(-1), out)
    assert rc == 0
    return complex(out[0], out[1])


def frozen_fd(verts64, faces, pj, ij, ids, cnt, w, coords, h, richardson=False):
    """Central differences of sum_j w_j sigma_j along frozen paths for
    coordinates coords (3*vertex + axis); with richardson, the O(h^4)
    combination of steps h and h/2. Returns (grad, sigma)."""
    v = np.ascontiguousarray(verts64, np.float64).reshape(-1)
    f = np.ascontiguousarray(faces, np.uint32).reshape(-1)
    pj = np.ascontiguousarray(pj, np.float64)
    ij = np.ascontiguousarray(ij, np.int32)
    nj = len(pj)
    coords = np.ascontiguousarray(coords, np.int32)
    g = np.zeros(len(coords), np.float64)
    sg = np.zeros(nj, np.float64)
    rc = lib().orc_frozen_fd(v, len(v) // 3, f, len(f) // 3, nj, pj.reshape(-1), ij.reshape(-1),
                             np.ascontiguousarray(ids, np.int32).reshape(-1),
                             np.ascontiguousarray(cnt, np.int32).reshape(-1), np.ascontiguousarray(w, np.float64),
                             len(coords), coords, float(h), int(richardson), g, sg)
    assert rc == 0, rc
    return g, sg


# ------------------------------------------- reference HIP kernels (GPU)
REF_GPU_SO = {
    "f32": os.path.join(ROOT, "oracle", "_ref", "libref_gpu_f32.so"),
    "f64": os.path.join(ROOT, "oracle", "_ref", "libref_gpu_f64.so"),
}
REF_LOADER_SO = os.path.join(ROOT, "oracle", "_ref", "libref_loader.so")
_ref_loader = None


def ref_loader():
    """The reference's own model loader (src/modelLoader.cpp with its
    vendored tinyobjloader / tinygltf) built by oracle/Makefile; None when not
    built."""
    global _ref_loader
    if _ref_loader is None and os.path.exists(REF_LOADER_SO):
        L = C.CDLL(REF_LOADER_SO)
        L.refload_open.restype = _P
        L.refload_open.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        L.refload_info.argtypes = [_P, _lp]
        L.refload_copy.argtypes = [_P, _fp, _up]
        L.refload_close.argtypes = [_P]
        _ref_loader = L
    return _ref_loader


def ref_load_model(path):
    """(float32 positions (V, 3), uint32 faces (F, 3)) from the reference
    loader; raises with its message when it fails."""
    L = ref_loader()
    err = C.create_string_buffer(4096)
    h = L.refload_open(os.fsencode(path), err, 4096)
    if not h:
        raise ValueError(err.value.decode(errors="replace"))
    try:
        n = np.zeros(2, np.int64)
        L.refload_info(h, n)
        v = np.zeros(int(n[0]) * 3, np.float32)
        f = np.zeros(int(n[1]) * 3, np.uint32)
        L.refload_copy(h, v, f)
    finally:
        L.refload_close(h)
    return v.reshape(-1, 3), f.reshape(-1, 3)


_ref_gpu = {}


def ref_gpu_lib(prec="f64"):
    """The reference's own rtKernels.cu / poKernels.cu / bvhBuilder.cpp built
    from its sources by oracle/Makefile, with oracle/ref_gpu_harness.hip as
    the launcher; None when not built."""
    if prec not in _ref_gpu:
        if not os.path.exists(REF_GPU_SO[prec]):
            return None
        L = C.CDLL(REF_GPU_SO[prec])
        L.refgpu_real_size.restype = C.c_int
        L.refgpu_create.restype = _P
        L.refgpu_create.argtypes = [_fp, C.c_int, _up, C.c_int, C.c_double]
        L.refgpu_destroy.argtypes = [_P]
        L.refgpu_sweep.argtypes = [_P, _dp, _ip, _ip, C.c_int, C.c_int, _dp, _dp]
        L.refgpu_sweep.restype = C.c_int
        L.refgpu_rays.argtypes = [_P, _dp, _ip, C.c_int, _ip, _dp]
        L.refgpu_rays.restype = C.c_int
        _ref_gpu[prec] = L
    return _ref_gpu[prec]


class RefGpu:
    """Reference HIP kernels on one mesh (reference BVH, reference launch)."""

    def __init__(self, verts, faces, prec="f64", scale=1.0):
        self.L = ref_gpu_lib(prec)
        if self.L is None:
            raise RuntimeError(f"reference GPU library not built: {REF_GPU_SO[prec]}")
        v = np.ascontiguousarray(verts, np.float32).reshape(-1)
        f = np.ascontiguousarray(faces, np.uint32).reshape(-1)
        self.h = self.L.refgpu_create(v, len(v) // 3, f, len(f) // 3, float(scale))
        if not self.h:
            raise RuntimeError("refgpu_create failed")

    def close(self):
        if getattr(self, "h", None):
            self.L.refgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @staticmethod
    def _pp(p: Params):
        pd = np.array([p.theta_start, p.theta_end, p.phi_start, p.phi_end, p.grid_size, p.freq, p.reflection_const],
                      np.float64)
        pi = np.array([p.theta_samples, p.phi_samples, p.n_x, p.n_y, p.max_bounces], np.int32)
        return pd, pi

    def sweep(self, p: Params, angles=None, host_sum=False):
        """(complex sum per angle, (rt_ms, po_ms) kernel totals). host_sum:
        the sums are the reference PO term of the RT kernel's per-ray outputs
        added on the host, without the PO kernel's block reduction."""
        a = np.arange(p.n_angles, dtype=np.int32) if angles is None else np.ascontiguousarray(angles, np.int32)
        pd, pi = self._pp(p)
        reim = np.zeros(2 * len(a), np.float64)
        t = np.zeros(2, np.float64)
        rc = self.L.refgpu_sweep(self.h, pd, pi, a, len(a), int(host_sum), reim, t)
        if rc:
            raise RuntimeError(f"refgpu_sweep failed ({rc})")
        return reim[0::2] + 1j * reim[1::2], (float(t[0]), float(t[1]))

    def rays(self, p: Params, angle):
        pd, pi = self._pp(p)
        n = p.n_x * p.n_y
        cnt = np.zeros(n, np.int32)
        dist = np.zeros(n, np.float64)
        rc = self.L.refgpu_rays(self.h, pd, pi, int(angle), cnt, dist)
        if rc:
            raise RuntimeError(f"refgpu_rays failed ({rc})")
        return cnt, dist
