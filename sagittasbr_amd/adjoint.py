"""Differentiable monostatic RCS: torch.autograd binding of libsbr's
frozen-path adjoint (sbr_adjoint in include/sbr.h).

    sigma = monostatic_rcs(ctx, vertices, faces, theta_deg, phi_deg, freq,
                           grid_size=2.5, n_x=250, n_y=250, max_bounces=4)
    (w * sigma).sum().backward()      # -> vertices.grad, freq.grad

The forward pass is one batched mixed-precision sweep (each ray ends with
the FP64 reference outcome); the backward pass is one sbr_adjoint call, which
records the same paths (hit faces frozen) and differentiates sigma along
them, recomputing the bounce geometry in FP64 from the recorded faces. Its
d sigma / d f comes with the vertex gradient, so nothing is computed twice.
Vertices are uploaded as float32, the precision of the reference's mesh
positions (src/app/rtRcsEntry.cu:136-144), so gradients are with respect to
those rounded coordinates.

Meshes (BVH build and upload) are cached per context by the bytes of their
float32 vertices and faces: repeated evaluations of one geometry reuse its
mesh, and at most MESH_CACHE meshes stay alive (the oldest is destroyed; a
backward pass whose mesh was destroyed builds it again). Everything is freed
with the context.

torch is used only to carry tensors in and out; the computation is the HIP
library.
"""
from __future__ import annotations

import hashlib
from collections import OrderedDict

import numpy as np
import torch

from . import _native as N

MESH_CACHE = 2


def _mesh_for(engine, v32: np.ndarray, f32: np.ndarray) -> int:
    """The engine's mesh for these vertices and faces, built on first use."""
    cache = engine.__dict__.setdefault("_adjoint_meshes", OrderedDict())
    key = hashlib.blake2b(v32.tobytes() + b"|" + f32.tobytes(), digest_size=20).digest()
    mid = cache.get(key)
    if mid is not None:
        cache.move_to_end(key)
        return mid
    mid = engine.mesh(v32, f32)
    cache[key] = mid
    while len(cache) > MESH_CACHE:
        _, old = cache.popitem(last=False)
        engine.destroy_mesh(old)
    return mid


def _jobs(mesh, theta_deg, phi_deg, freq, grid_size, n_x, n_y, max_bounces):
    return N.direction_jobs(mesh, theta_deg, phi_deg, float(freq), grid_size, n_x, n_y, max_bounces)


class MonostaticRCS(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vertices, freq, faces, theta_deg, phi_deg, grid_size, n_x, n_y, max_bounces, engine):
        v = np.ascontiguousarray(vertices.detach().cpu().numpy(), np.float32).reshape(-1, 3)
        f = np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
        mesh = _mesh_for(engine, v, f)
        plan = engine.plan(_jobs(mesh, theta_deg, phi_deg, freq, grid_size, n_x, n_y, max_bounces), N.SBR_MIXED,
                           N.SBR_BATCHED)
        try:
            plan.run()
            plan.wait()
            _, sigma, _ = plan.results()
        finally:
            plan.close()
        ctx.engine = engine
        ctx.v, ctx.f = v, f
        ctx.args = (theta_deg, phi_deg, float(freq), grid_size, n_x, n_y, max_bounces)
        ctx.vshape, ctx.vdtype, ctx.vdevice = vertices.shape, vertices.dtype, vertices.device
        ctx.fdtype = freq.dtype if torch.is_tensor(freq) else torch.float64
        ctx.fdevice = freq.device if torch.is_tensor(freq) else torch.device("cpu")
        return torch.from_numpy(sigma).to(vertices.device)

    @staticmethod
    def backward(ctx, grad_sigma):
        gv = gf = None
        if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
            w = grad_sigma.detach().cpu().numpy().astype(np.float64)
            mesh = _mesh_for(ctx.engine, ctx.v, ctx.f)
            _, dxyz, dsdf = ctx.engine.adjoint(mesh, _jobs(mesh, *ctx.args), w)
            if ctx.needs_input_grad[0]:
                gv = torch.from_numpy(dxyz).to(device=ctx.vdevice, dtype=ctx.vdtype).reshape(ctx.vshape)
            if ctx.needs_input_grad[1]:
                gf = torch.tensor(float(np.dot(w, dsdf)), dtype=ctx.fdtype, device=ctx.fdevice)
        return gv, gf, None, None, None, None, None, None, None, None


def monostatic_rcs(engine, vertices, faces, theta_deg, phi_deg, freq, grid_size=2.5, n_x=250, n_y=250,
                   max_bounces=4):
    """sigma [m^2] per direction, differentiable in `vertices` (n_verts x 3)
    and `freq` (a scalar tensor or float, Hz)."""
    if not torch.is_tensor(freq):
        freq = torch.tensor(float(freq), dtype=torch.float64)
    return MonostaticRCS.apply(vertices, freq, faces, np.asarray(theta_deg, np.float64),
                               np.asarray(phi_deg, np.float64), grid_size, n_x, n_y, max_bounces, engine)
