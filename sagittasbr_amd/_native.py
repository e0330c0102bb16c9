"""ctypes binding of libsbr.so (C ABI in include/sbr.h).

The shared library is built in-tree by ``make -C sagittasbr_amd`` (or
``__graft_entry__.build()``). Loading fails loudly when it is missing: there
is no CPU fallback for the ray path.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsbr.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "sbr.h")

SBR_FP32, SBR_FP64, SBR_MIXED = 0, 1, 2
SBR_PER_ANGLE, SBR_BATCHED = 0, 1
PRECISIONS = {"fp32": SBR_FP32, "fp64": SBR_FP64, "mixed": SBR_MIXED}
METHODS = {"per_angle": SBR_PER_ANGLE, "batched": SBR_BATCHED}


class SbrError(RuntimeError):
    pass


class BvhOptions(C.Structure):
    """bvhBuildOptions (reference include/scene/bvhBuilder.hpp:13-20)."""

    _fields_ = [
        ("leaf_size", C.c_int),
        ("bin_count", C.c_int),
        ("max_parallel_depth", C.c_int),
        ("min_triangles_for_parallel", C.c_int),
        ("enable_parallel", C.c_int),
        ("algorithm", C.c_int),
    ]

    @classmethod
    def default(cls, **kw):
        o = cls(4, 16, 6, 4096, 1, 1)
        for k, v in kw.items():
            setattr(o, k, v)
        return o


class SweepParams(C.Structure):
    """simulationConfig (reference include/app/config.hpp:7-27)."""

    _fields_ = [
        ("phi_start", C.c_double),
        ("phi_end", C.c_double),
        ("phi_samples", C.c_int),
        ("theta_start", C.c_double),
        ("theta_end", C.c_double),
        ("theta_samples", C.c_int),
        ("grid_size", C.c_double),
        ("n_x", C.c_int),
        ("n_y", C.c_int),
        ("max_bounces", C.c_int),
        ("reflection_const", C.c_double),
        ("freq", C.c_double),
    ]

    @classmethod
    def make(cls, **kw):
        # reference defaults, include/app/config.hpp:8-25
        d = dict(phi_start=0.0, phi_end=180.0, phi_samples=181, theta_start=90.0, theta_end=90.0, theta_samples=1,
                 grid_size=10.0, n_x=5000, n_y=5000, max_bounces=20, reflection_const=1.0, freq=10.0e9)
        d.update(kw)
        return cls(**d)

    @property
    def n_angles(self):
        return self.theta_samples * self.phi_samples


class Job(C.Structure):
    _fields_ = [
        ("mesh", C.c_int),
        ("max_bounces", C.c_int),
        ("n_x", C.c_int),
        ("n_y", C.c_int),
        ("freq", C.c_double),
        ("theta_deg", C.c_double),
        ("phi_deg", C.c_double),
        ("grid_size", C.c_double),
        ("reflection_const", C.c_double),
    ]


ABI_VERSION = 5  # SBR_ABI_VERSION of the structures below


class Tuning(C.Structure):
    _fields_ = [("wg_per_cu", C.c_int), ("lds_stack", C.c_int), ("tile_rays", C.c_int), ("use_graph", C.c_int),
                ("ray_pool", C.c_int), ("traversal", C.c_int), ("top_levels", C.c_int), ("node_width", C.c_int),
                ("cull", C.c_int)]


class MeshInfo(C.Structure):
    _fields_ = [
        ("n_triangles", C.c_int),
        ("n_nodes", C.c_int),
        ("n_gpu_nodes", C.c_int),
        ("depth", C.c_int),
        ("bounds", C.c_double * 6),
        ("build_ms", C.c_double),
    ]


class Counters(C.Structure):
    _fields_ = [
        ("rays", C.c_uint64),
        ("node_visits", C.c_uint64),
        ("tri_tests", C.c_uint64),
        ("ray_bounces", C.c_uint64),
        ("queries", C.c_uint64),
        ("retraced", C.c_uint64),
        ("fp64_queries", C.c_uint64),
    ]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


_lib = None
_P = C.c_void_p


def header_functions():
    """Names of the functions declared in include/sbr.h."""
    txt = open(HEADER_PATH).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(sbr_\w+)\s*\(", txt, flags=re.M)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    path = os.environ.get("SAGITTASBR_LIB", LIB_PATH)  # another build of libsbr.so
    if not os.path.exists(path):
        raise SbrError(f"libsbr.so not built at {path}; run `make -C sagittasbr_amd` first")
    L = C.CDLL(path)
    i32p = C.POINTER(C.c_int32)
    f64p = C.POINTER(C.c_double)
    sig = {
        "sbr_last_error": (C.c_char_p, [_P]),
        "sbr_abi_version": (C.c_int, []),
        "sbr_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
        "sbr_ctx_destroy": (C.c_int, [_P]),
        "sbr_ctx_cu_count": (C.c_int, [_P, C.POINTER(C.c_int)]),
        "sbr_ctx_set_stream": (C.c_int, [_P, _P]),
        "sbr_mesh_create": (C.c_int, [_P, _P, C.c_int, _P, C.c_int, C.c_double, _P, C.POINTER(C.c_int)]),
        "sbr_mesh_info_get": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(MeshInfo)]),
        "sbr_mesh_permutation": (C.c_int, [_P, C.c_int, C.c_int, _P]),
        "sbr_mesh_export_bvh": (C.c_int, [_P, C.c_int, C.c_int, _P, _P, C.POINTER(C.c_int)]),
        "sbr_bvh_build": (C.c_int, [_P, C.c_int, _P, C.c_int, C.c_double, _P, C.c_int, C.POINTER(_P)]),
        "sbr_bvh_info": (C.c_int, [_P, _P]),
        "sbr_bvh_export": (C.c_int, [_P, _P, _P, _P, _P]),
        "sbr_bvh_destroy": (C.c_int, [_P]),
        "sbr_model_load": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
        "sbr_model_info": (C.c_int, [_P, _P]),
        "sbr_model_copy": (C.c_int, [_P, _P, _P]),
        "sbr_model_destroy": (C.c_int, [_P]),
        "sbr_obj_parse_number": (C.c_int, [C.c_char_p, f64p]),
        "sbr_angle_frame": (C.c_int, [C.POINTER(SweepParams), C.c_int, C.c_int, _P, _P]),
        "sbr_sweep_jobs": (C.c_int, [C.POINTER(SweepParams), C.c_int, C.c_int, _P, C.c_int]),
        "sbr_plan_create": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, C.POINTER(_P)]),
        "sbr_plan_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
        "sbr_adjoint": (C.c_int, [_P, C.c_int, _P, C.c_int, _P, _P, _P, _P]),
        "sbr_mesh_destroy": (C.c_int, [_P, C.c_int]),
        "sbr_plan_autotune": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(Tuning), C.POINTER(C.c_double),
                                        C.POINTER(C.c_int)]),
        "sbr_plan_run": (C.c_int, [_P]),
        "sbr_plan_wait": (C.c_int, [_P]),
        "sbr_plan_last_ms": (C.c_int, [_P, f64p]),
        "sbr_plan_results": (C.c_int, [_P, _P, _P, _P]),
        "sbr_plan_count": (C.c_int, [_P, C.POINTER(Counters)]),
        "sbr_plan_count_traced": (C.c_int, [_P, C.POINTER(Counters)]),
        "sbr_plan_destroy": (C.c_int, [_P]),
        "sbr_sweep": (C.c_int, [_P, C.c_int, C.POINTER(SweepParams), C.c_int, C.c_int, _P, _P, _P]),
        "sbr_sweep_noalloc": (C.c_int, [_P, C.c_int, C.POINTER(SweepParams), C.c_int, _P, _P, _P]),
        "sbr_trace_rays": (C.c_int, [_P, C.c_int, C.POINTER(SweepParams), C.c_int, C.c_int, _P, _P, _P, _P, _P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if L.sbr_abi_version() != ABI_VERSION:
        raise SbrError(f"{path} has ABI {L.sbr_abi_version()}, this binding expects {ABI_VERSION}: rebuild it")
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data


def _check(rc, ctx=None):
    if rc != 0:
        msg = lib().sbr_last_error(ctx)
        raise SbrError(msg.decode() if msg else "libsbr call failed")


def _mesh_arrays(verts, faces):
    v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
    f = np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
    return v, f


def host_bvh(verts, faces, precision=SBR_FP32, scale=1.0, options=None):
    """Build the BVH on the host only; returns dict of arrays."""
    v, f = _mesh_arrays(verts, faces)
    h = _P()
    opts = C.byref(options) if options is not None else None
    _check(lib().sbr_bvh_build(_ptr(v), len(v), _ptr(f), len(f), scale, opts, precision, C.byref(h)))
    try:
        info = np.zeros(4, np.int32)
        _check(lib().sbr_bvh_info(h, _ptr(info)))
        nt, nn, root, depth = (int(x) for x in info)
        tris = np.zeros((nt, 9), np.float64)
        perm = np.zeros(nt, np.int32)
        bounds = np.zeros((nn, 6), np.float64)
        links = np.zeros((nn, 4), np.int32)
        _check(lib().sbr_bvh_export(h, _ptr(tris), _ptr(perm), _ptr(bounds), _ptr(links)))
    finally:
        lib().sbr_bvh_destroy(h)
    return dict(tris=tris, perm=perm, bounds=bounds, links=links, root=root, depth=depth)


def load_model(path):
    """Load .obj / .gltf / .glb like the reference's modelLoader; returns
    (verts float32 (V,3), faces uint32 (F,3))."""
    h = _P()
    _check(lib().sbr_model_load(os.fsencode(path), C.byref(h)))
    try:
        info = np.zeros(2, np.int64)
        _check(lib().sbr_model_info(h, _ptr(info)))
        v = np.zeros((int(info[0]), 3), np.float32)
        f = np.zeros((int(info[1]), 3), np.uint32)
        _check(lib().sbr_model_copy(h, _ptr(v), _ptr(f)))
    finally:
        lib().sbr_model_destroy(h)
    return v, f


def obj_parse_number(token: str):
    """tinyobj's number parsing of one token (None when it fails)."""
    out = C.c_double()
    if lib().sbr_obj_parse_number(token.encode(), C.byref(out)) != 0:
        return None
    return out.value


def angle_frame(params: SweepParams, angle: int, precision=SBR_FP32):
    out = np.zeros(12, np.float64)
    deg = np.zeros(2, np.float64)
    _check(lib().sbr_angle_frame(C.byref(params), angle, precision, _ptr(out), _ptr(deg)))
    return out.reshape(4, 3), deg


def sweep_jobs(params: SweepParams, mesh: int, precision=SBR_FP32):
    n = params.n_angles
    jobs = (Job * n)()
    _check(lib().sbr_sweep_jobs(C.byref(params), mesh, precision, C.cast(jobs, _P), n))
    return jobs


def direction_jobs(mesh: int, theta_deg, phi_deg, freq, grid_size, n_x, n_y, max_bounces, reflection_const=1.0):
    """Jobs for explicit (theta, phi) directions in degrees (one job each)."""
    th = np.atleast_1d(np.asarray(theta_deg, np.float64))
    ph = np.atleast_1d(np.asarray(phi_deg, np.float64))
    if th.shape != ph.shape:
        raise ValueError("theta and phi must have the same length")
    jobs = (Job * len(th))()
    for k in range(len(th)):
        jobs[k] = Job(mesh, max_bounces, n_x, n_y, float(freq), float(th[k]), float(ph[k]), float(grid_size),
                      float(reflection_const))
    return jobs


class Plan:
    def __init__(self, ctx: "Context", jobs, precision=SBR_FP32, method=SBR_BATCHED):
        self.ctx = ctx
        if not isinstance(jobs, C.Array):
            arr = (Job * len(jobs))()
            for i, j in enumerate(jobs):
                arr[i] = j
            jobs = arr
        self.n = len(jobs)
        self.precision = precision
        self.h = _P()
        _check(lib().sbr_plan_create(ctx.h, C.cast(jobs, _P), self.n, precision, method, C.byref(self.h)), ctx.h)

    def tune(self, wg_per_cu=0, lds_stack=0, tile_rays=0, use_graph=0, ray_pool=0, traversal=0, top_levels=0,
             node_width=0, cull=0):
        t = Tuning(wg_per_cu, lds_stack, tile_rays, use_graph, ray_pool, traversal, top_levels, node_width, cull)
        _check(lib().sbr_plan_set_tuning(self.h, C.byref(t)), self.ctx.h)

    def autotune(self, sample_jobs=0, repeats=1):
        """Time the plan over the tuning space, apply and return the fastest
        (results are independent of the tuning)."""
        t, ms, n = Tuning(), C.c_double(), C.c_int()
        _check(lib().sbr_plan_autotune(self.h, sample_jobs, repeats, C.byref(t), C.byref(ms), C.byref(n)),
               self.ctx.h)
        return {"wg_per_cu": t.wg_per_cu, "lds_stack": t.lds_stack, "tile_rays": t.tile_rays,
                "use_graph": t.use_graph, "ray_pool": t.ray_pool, "traversal": t.traversal,
                "top_levels": t.top_levels, "node_width": t.node_width, "cull": t.cull, "best_ms": ms.value,
                "candidates": n.value}

    def run(self):
        _check(lib().sbr_plan_run(self.h), self.ctx.h)

    def wait(self):
        _check(lib().sbr_plan_wait(self.h), self.ctx.h)

    def last_ms(self):
        ms = C.c_double()
        _check(lib().sbr_plan_last_ms(self.h, C.byref(ms)), self.ctx.h)
        return ms.value

    def results(self):
        reim = np.zeros(2 * self.n, np.float64)
        sig = np.zeros(self.n, np.float64)
        db = np.zeros(self.n, np.float64)
        _check(lib().sbr_plan_results(self.h, _ptr(reim), _ptr(sig), _ptr(db)), self.ctx.h)
        return reim[0::2] + 1j * reim[1::2], sig, db

    def count(self, traced=False):
        """Work counters of one run: over every ray of each job's culled
        rectangle, or (traced=True) over the rays a timed run traces."""
        c = Counters()
        fn = lib().sbr_plan_count_traced if traced else lib().sbr_plan_count
        _check(fn(self.h, C.byref(c)), self.ctx.h)
        return c.as_dict()

    def close(self):
        if getattr(self, "h", None):
            lib().sbr_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Context:
    """A libsbr context bound to one HIP device."""

    def __init__(self, device=0):
        self.h = _P()
        self._mesh_nverts = {}
        _check(lib().sbr_ctx_create(device, C.byref(self.h)))

    def close(self):
        if getattr(self, "h", None):
            lib().sbr_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_stream(self, stream=None):
        """Issue the context's work on a caller's HIP stream (a raw hipStream_t
        handle from the HIP runtime libsbr.so links, /opt/rocm/lib), or on its
        own with None. torch's ROCm wheel carries a HIP runtime of its own:
        its stream handles are not valid here."""
        _check(lib().sbr_ctx_set_stream(self.h, None if stream is None else C.c_void_p(int(stream))), self.h)

    @property
    def cu_count(self):
        n = C.c_int()
        _check(lib().sbr_ctx_cu_count(self.h, C.byref(n)), self.h)
        return n.value

    def mesh(self, verts, faces, scale=1.0, options=None) -> int:
        v, f = _mesh_arrays(verts, faces)
        mid = C.c_int()
        opts = C.byref(options) if options is not None else None
        _check(lib().sbr_mesh_create(self.h, _ptr(v), len(v), _ptr(f), len(f), scale, opts, C.byref(mid)), self.h)
        self._mesh_nverts[mid.value] = len(v)
        return mid.value

    def destroy_mesh(self, mesh):
        _check(lib().sbr_mesh_destroy(self.h, mesh), self.h)
        self._mesh_nverts.pop(mesh, None)

    def mesh_info(self, mesh, precision=SBR_FP32):
        mi = MeshInfo()
        _check(lib().sbr_mesh_info_get(self.h, mesh, precision, C.byref(mi)), self.h)
        return dict(n_triangles=mi.n_triangles, n_nodes=mi.n_nodes, n_gpu_nodes=mi.n_gpu_nodes, depth=mi.depth,
                    bounds=list(mi.bounds), build_ms=mi.build_ms)

    def permutation(self, mesh, precision=SBR_FP32):
        n = self.mesh_info(mesh, precision)["n_triangles"]
        p = np.zeros(n, np.int32)
        _check(lib().sbr_mesh_permutation(self.h, mesh, precision, _ptr(p)), self.h)
        return p

    def export_bvh(self, mesh, precision=SBR_FP32):
        n = self.mesh_info(mesh, precision)["n_nodes"]
        b = np.zeros((n, 6), np.float64)
        l = np.zeros((n, 4), np.int32)
        r = C.c_int()
        _check(lib().sbr_mesh_export_bvh(self.h, mesh, precision, _ptr(b), _ptr(l), C.byref(r)), self.h)
        return b, l, r.value

    def sweep(self, mesh, params: SweepParams, precision=SBR_FP32, method=SBR_BATCHED):
        n = params.n_angles
        reim = np.zeros(2 * n, np.float64)
        sig = np.zeros(n, np.float64)
        db = np.zeros(n, np.float64)
        _check(lib().sbr_sweep(self.h, mesh, C.byref(params), precision, method, _ptr(reim), _ptr(sig), _ptr(db)),
               self.h)
        return reim[0::2] + 1j * reim[1::2], sig, db

    def sweep_noalloc(self, mesh, params: SweepParams, precision=SBR_FP32):
        """Batched sweep through the allocation-free entry point (its device
        storage is a fixed array; see sbr_sweep_noalloc in include/sbr.h)."""
        n = params.n_angles
        reim = np.zeros(2 * n, np.float64)
        sig = np.zeros(n, np.float64)
        db = np.zeros(n, np.float64)
        _check(lib().sbr_sweep_noalloc(self.h, mesh, C.byref(params), precision, _ptr(reim), _ptr(sig), _ptr(db)),
               self.h)
        return reim[0::2] + 1j * reim[1::2], sig, db

    def adjoint(self, mesh, jobs, dloss_dsigma=None):
        """Frozen-path adjoint: returns (sigma[n_jobs], dloss_dxyz[n_verts, 3],
        dsigma_dfreq[n_jobs]); dloss_dxyz is with respect to the coordinates
        given to `mesh` (before its scale)."""
        if not isinstance(jobs, C.Array):
            arr = (Job * len(jobs))()
            for i, j in enumerate(jobs):
                arr[i] = j
            jobs = arr
        n = len(jobs)
        nv = self._mesh_nverts[mesh]
        w = np.zeros(n, np.float64) if dloss_dsigma is None else np.ascontiguousarray(dloss_dsigma, np.float64)
        if w.shape != (n,):
            raise ValueError("dloss_dsigma must have one value per job")
        sig = np.zeros(n, np.float64)
        dx = np.zeros(nv * 3, np.float64)
        df = np.zeros(n, np.float64)
        _check(lib().sbr_adjoint(self.h, mesh, C.cast(jobs, _P), n, _ptr(w), _ptr(sig), _ptr(dx), _ptr(df)), self.h)
        return sig, dx.reshape(nv, 3), df

    def plan(self, jobs, precision=SBR_FP32, method=SBR_BATCHED) -> Plan:
        return Plan(self, jobs, precision, method)

    def trace_rays(self, mesh, params: SweepParams, angle, precision=SBR_FP32):
        nr = params.n_x * params.n_y
        mb = max(params.max_bounces, 1)
        ids = np.full(nr * mb, -1, np.int32)
        cnt = np.zeros(nr, np.int32)
        L = np.zeros(nr, np.float64)
        reim = np.zeros(2 * nr, np.float64)
        flag = np.zeros(nr, np.int8)
        _check(lib().sbr_trace_rays(self.h, mesh, C.byref(params), angle, precision, _ptr(ids), _ptr(cnt), _ptr(L),
                                    _ptr(reim), _ptr(flag)), self.h)
        return dict(ids=ids.reshape(nr, mb)[:, : params.max_bounces], count=cnt, L=L, reim=reim.reshape(nr, 2),
                    retraced=flag)
