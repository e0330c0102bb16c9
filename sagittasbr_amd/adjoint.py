"""Differentiable monostatic RCS: torch.autograd binding of libsbr's
frozen-path adjoint (sbr_adjoint in include/sbr.h).

    sigma = monostatic_rcs(ctx, vertices, faces, theta_deg, phi_deg, freq,
                           grid_size=2.5, n_x=250, n_y=250, max_bounces=4)
    (w * sigma).sum().backward()      # -> vertices.grad, freq.grad

The forward pass is the mixed-precision engine (each ray ends with the FP64
reference outcome); the backward pass differentiates sigma along the paths
the forward pass found (hit faces frozen), recomputing the bounce geometry
in FP64 from the recorded faces. Vertices are uploaded as float32, the
precision of the reference's mesh positions (src/app/rtRcsEntry.cu:136-144),
so gradients are with respect to those rounded coordinates.

torch is used only to carry tensors in and out; the computation is the HIP
library.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as N


class MonostaticRCS(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vertices, freq, faces, theta_deg, phi_deg, grid_size, n_x, n_y, max_bounces, engine):
        v = vertices.detach().cpu().numpy().astype(np.float32)
        f = np.ascontiguousarray(faces, np.uint32)
        mesh = engine.mesh(v, f)
        try:
            jobs = N.direction_jobs(mesh, theta_deg, phi_deg, float(freq), grid_size, n_x, n_y, max_bounces)
            sigma, _, dsdf = engine.adjoint(mesh, jobs)
        except Exception:
            engine.destroy_mesh(mesh)
            raise
        ctx.engine = engine
        ctx.mesh = mesh
        ctx.jobs = jobs
        ctx.dsdf = dsdf
        ctx.vshape = vertices.shape
        ctx.vdtype = vertices.dtype
        ctx.vdevice = vertices.device
        ctx.fdtype = freq.dtype if torch.is_tensor(freq) else torch.float64
        ctx.fdevice = freq.device if torch.is_tensor(freq) else torch.device("cpu")
        if not (ctx.needs_input_grad[0] or ctx.needs_input_grad[1]):
            engine.destroy_mesh(mesh)
            ctx.mesh = None
        return torch.from_numpy(sigma).to(vertices.device)

    @staticmethod
    def backward(ctx, grad_sigma):
        w = grad_sigma.detach().cpu().numpy().astype(np.float64)
        gv = gf = None
        if ctx.mesh is not None:
            try:
                if ctx.needs_input_grad[0]:
                    _, dxyz, _ = ctx.engine.adjoint(ctx.mesh, ctx.jobs, w)
                    gv = torch.from_numpy(dxyz).to(device=ctx.vdevice, dtype=ctx.vdtype).reshape(ctx.vshape)
            finally:
                ctx.engine.destroy_mesh(ctx.mesh)
                ctx.mesh = None
        if ctx.needs_input_grad[1]:
            gf = torch.tensor(float(np.dot(w, ctx.dsdf)), dtype=ctx.fdtype, device=ctx.fdevice)
        return gv, gf, None, None, None, None, None, None, None, None


def monostatic_rcs(engine, vertices, faces, theta_deg, phi_deg, freq, grid_size=2.5, n_x=250, n_y=250,
                   max_bounces=4):
    """sigma [m^2] per direction, differentiable in `vertices` (n_verts x 3)
    and `freq` (a scalar tensor or float, Hz)."""
    if not torch.is_tensor(freq):
        freq = torch.tensor(float(freq), dtype=torch.float64)
    return MonostaticRCS.apply(vertices, freq, faces, np.asarray(theta_deg, np.float64),
                               np.asarray(phi_deg, np.float64), grid_size, n_x, n_y, max_bounces, engine)
