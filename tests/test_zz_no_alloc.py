"""The allocation-free sweep (sbr_sweep_noalloc) against sbr_sweep.

sbr_sweep_noalloc runs the batched sweep out of a fixed device array instead
of buffers the library allocates. Its per-job sums must equal sbr_sweep's to
the bit (both reduce the same per-pass partials in the same fixed order), on
two calls in a row, the second smaller than the first (the array keeps the
first call's contents), in every precision. A sweep too large for the array
is refused on the host before anything is launched.
"""
import numpy as np
import pytest

S = pytest.importorskip("sagittasbr_amd")
from sagittasbr_amd import meshes as M  # noqa: E402

pytestmark = pytest.mark.gpu


def _params(**kw):
    d = dict(theta_start=60, theta_end=120, theta_samples=4, phi_start=0, phi_end=350, phi_samples=9, grid_size=2.5,
             n_x=160, n_y=160, max_bounces=5, freq=3e9)
    d.update(kw)
    return S.SweepParams.make(**d)


def test_noalloc_matches_sweep_two_calls():
    ctx = S.Context(0)
    try:
        v, f = M.icosphere(4)
        mesh = ctx.mesh(v, f)
        big = _params()
        small = _params(theta_samples=2, phi_samples=3, n_x=64, n_y=64, phi_end=200)
        # FP64 and mixed need the FP64 tree, which the allocation-free call
        # does not build
        for prec in (S.SBR_FP64, S.SBR_MIXED):
            print(f"  [step] refused without an FP64 tree, precision {prec}", flush=True)
            with pytest.raises(S.SbrError, match="FP64 tree"):
                ctx.sweep_noalloc(mesh, small, prec)
        ctx.mesh_info(mesh, S.SBR_FP64)  # builds the FP64 tree (allocating, at set-up)
        for prec in (S.SBR_FP32, S.SBR_FP64, S.SBR_MIXED):
            for name, p in (("first", big), ("second, smaller", small)):
                print(f"  [step] precision {prec}, {name} call: {p.n_angles} angles, {p.n_x}x{p.n_y} rays",
                      flush=True)
                z0, s0, d0 = ctx.sweep(mesh, p, prec, S.SBR_BATCHED)
                z1, s1, d1 = ctx.sweep_noalloc(mesh, p, prec)
                assert np.array_equal(z1, z0)
                assert np.array_equal(s1, s0)
                assert np.array_equal(d1, d0)
                assert np.all(s0 > 0)
    finally:
        ctx.close()


def test_noalloc_refuses_oversize_sweep():
    ctx = S.Context(0)
    try:
        v, f = M.icosphere(2)
        mesh = ctx.mesh(v, f)
        # 8 angles of 30000 x 30000 rays: ~18 MB of FP32 pass partials per
        # angle after culling to the sphere's silhouette
        p = _params(theta_samples=2, phi_samples=4, n_x=30000, n_y=30000)
        print("  [step] oversize sweep refused", flush=True)
        with pytest.raises(S.SbrError, match="bytes of device scratch"):
            ctx.sweep_noalloc(mesh, p, S.SBR_FP32)
        # the context stays usable
        print("  [step] small sweep after the refusal", flush=True)
        q = _params(theta_samples=1, phi_samples=2, n_x=48, n_y=48)
        z0, _, _ = ctx.sweep(mesh, q, S.SBR_FP32, S.SBR_BATCHED)
        z1, _, _ = ctx.sweep_noalloc(mesh, q, S.SBR_FP32)
        assert np.array_equal(z1, z0)
    finally:
        ctx.close()
