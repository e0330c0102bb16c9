"""GPU parity at the benchmark sizes (sphere: 360 angles of 250 x 250 rays;
trihedral+duct: 8,281 angles of 535 x 535 rays, 10 bounces; airframe:
3.2M triangles at 1, 3 and 10 GHz, every 31st of its 33,480 angles; 64
uneven bodies: all 23,040 jobs in one plan, eager and hipGraph replays).

At these sizes the oracle checks a seeded sample of angles, and the rest is
held by size-independent properties:
  * the FP64 kernel equals the FP64 oracle on the sampled angles (1e-9
    relative; sin/cos come from different libraries);
  * the mixed mode is within 0.1 dB of the FP64 kernel at every angle, and
    its amplitudes within 1e-9 of the largest (every ray ends as the FP64 ray
    ends; only the phase's sin/cos differ);
  * per-angle and batched launches give bit-identical sums, a repeated run
    gives bit-identical sums, and a job's sum does not depend on which other
    jobs share its plan.
"""
import numpy as np
import pytest

import oracle_lib as O

S = pytest.importorskip("sagittasbr_amd")
from sagittasbr_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu


def _step(msg):
    print(f"  [step] {msg}", flush=True)


def _db(z):
    return 10 * np.log10(np.maximum(4 * np.pi * np.abs(z) ** 2, 1e-10))


def _run(ctx, jobs, prec, method):
    pl = ctx.plan((S.Job * len(jobs))(*jobs), prec, method)
    try:
        pl.run()
        pl.wait()
        z, _, _ = pl.results()
    finally:
        pl.close()
    return z


def _mixed_close(zm, z64):
    assert np.max(np.abs(_db(zm) - _db(z64))) <= 0.1
    assert np.max(np.abs(zm - z64)) <= 1e-9 * np.abs(z64).max()


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_full_sweep(ctx, name):
    wl = getattr(W, name)()
    v, f = wl["meshes"][0]
    kw = wl["sweeps"][0][1]
    print(f"[full] {name}: {wl['desc']}", flush=True)
    mesh = ctx.mesh(v, f)
    p, op = S.SweepParams.make(**kw), O.Params(**kw)
    _step("fp64 batched, all angles")
    z64, _, _ = ctx.sweep(mesh, p, S.SBR_FP64, S.SBR_BATCHED)
    sample = np.sort(np.random.default_rng(11).choice(op.n_angles, 6, replace=False)).astype(np.int32)
    _step(f"fp64 oracle on angles {sample.tolist()}")
    ref, _ = O.Scene(v, f, 1).sweep(op, sample)
    np.testing.assert_allclose(z64[sample], ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
    _step("mixed batched, all angles")
    zm, _, _ = ctx.sweep(mesh, p, S.SBR_MIXED, S.SBR_BATCHED)
    _mixed_close(zm, z64)
    _step("mixed per-angle, all angles")
    zp, _, _ = ctx.sweep(mesh, p, S.SBR_MIXED, S.SBR_PER_ANGLE)
    assert np.array_equal(zm, zp)
    _step("mixed batched again")
    z2, _, _ = ctx.sweep(mesh, p, S.SBR_MIXED, S.SBR_BATCHED)
    assert np.array_equal(zm, z2)


def test_full_airframe_sample(ctx):
    wl = W.c3()
    v, f = wl["meshes"][0]
    print(f"[full] c3: {wl['desc']}", flush=True)
    mesh = ctx.mesh(v, f)

    def jobs(prec):
        out = []
        for _, kw in wl["sweeps"]:
            js = S.sweep_jobs(S.SweepParams.make(**kw), mesh, prec)
            out.extend(js[k] for k in range(len(js)))
        return out

    jm, j64 = jobs(S.SBR_MIXED)[::31], jobs(S.SBR_FP64)[::31]
    _step(f"fp64 batched, {len(j64)} jobs")
    z64 = _run(ctx, j64, S.SBR_FP64, S.SBR_BATCHED)
    _step(f"mixed batched, {len(jm)} jobs")
    zm = _run(ctx, jm, S.SBR_MIXED, S.SBR_BATCHED)
    _mixed_close(zm, z64)
    # every third of those jobs, alone in a plan of their own and in reverse
    # order, once batched and once per angle
    sub = jm[::3][::-1]
    _step(f"mixed batched, {len(sub)} of the jobs reversed")
    zs = _run(ctx, sub, S.SBR_MIXED, S.SBR_BATCHED)
    assert np.array_equal(zs, zm[::3][::-1])
    _step(f"mixed per-angle, {len(sub)} of the jobs reversed")
    zsp = _run(ctx, sub, S.SBR_MIXED, S.SBR_PER_ANGLE)
    assert np.array_equal(zsp, zs)


def test_full_uneven_batch(ctx):
    """64 bodies of 5k-400k triangles, each with its own frequency and grid,
    360 angles each: all 23,040 jobs in one plan, replayed from a hipGraph."""
    wl = W.c4()
    print(f"[full] c4: {wl['desc']}", flush=True)
    mids = [ctx.mesh(v, f) for v, f in wl["meshes"]]

    def jobs(prec):
        out = []
        for mi, kw in wl["sweeps"]:
            js = S.sweep_jobs(S.SweepParams.make(**kw), mids[mi], prec)
            out.extend(js[k] for k in range(len(js)))
        return out

    jm, j64 = jobs(S.SBR_MIXED), jobs(S.SBR_FP64)
    _step(f"fp64 batched, {len(j64)} jobs")
    z64 = _run(ctx, j64, S.SBR_FP64, S.SBR_BATCHED)
    # three seeded jobs against the FP64 oracle, each on its own body's scene
    for k in np.sort(np.random.default_rng(7).choice(len(j64), 3, replace=False)):
        mi, kw = wl["sweeps"][k // 360]
        v, f = wl["meshes"][mi]
        _step(f"fp64 oracle, job {k} (body {mi}, {len(f)} triangles, angle {k % 360})")
        ref, _ = O.Scene(v, f, 1).sweep(O.Params(**kw), np.array([k % 360], np.int32))
        body = z64[360 * (k // 360):360 * (k // 360 + 1)]
        np.testing.assert_allclose(z64[k], ref[0], rtol=1e-9, atol=1e-9 * np.abs(body).max())
    _step(f"mixed batched, {len(jm)} jobs, eager")
    zm = _run(ctx, jm, S.SBR_MIXED, S.SBR_BATCHED)
    _mixed_close(zm, z64)
    _step("mixed batched, hipGraph replays")
    pl = ctx.plan((S.Job * len(jm))(*jm), S.SBR_MIXED, S.SBR_BATCHED)
    try:
        pl.tune(0, 8, 32, 1, 2, 0)
        for _ in range(3):
            pl.run()
            pl.wait()
            zg, _, _ = pl.results()
            assert np.array_equal(zg, zm)
    finally:
        pl.close()
