"""GPU parity of the HIP ray path against the CPU oracle (through libsbr.so).

Bar:
  * FP32 kernel vs the oracle's FP32 twin (same fused-multiply-add places):
    per-ray hit ids, bounce counts and path lengths bit-exact.
  * FP64 kernel vs the FP64 oracle (the reference read literally in double):
    per-ray hit ids, bounce counts and path lengths bit-exact.
  * FP32 kernel vs FP64 oracle: every per-ray difference is an edge/grazing
    tie (FP64 outcome unstable under FP32-sized perturbations).
  * sigma: FP64 kernel within 1e-9 relative of the FP64 oracle (sin/cos
    come from different libraries); FP32 / mixed within 0.1 dB of FP64 on
    the sphere.
  * per-angle and batched methods give bit-identical sums; repeated runs
    are bit-identical.
"""
import numpy as np
import pytest

import oracle_lib as O
import parity_util as PU

S = pytest.importorskip("sagittasbr_amd")
from sagittasbr_amd import meshes as M  # noqa: E402

pytestmark = pytest.mark.gpu

C1 = dict(theta_start=90, theta_end=90, theta_samples=1, phi_start=0, phi_end=359, phi_samples=360, grid_size=2.5,
          n_x=250, n_y=250, max_bounces=5, freq=3e9)
C2S = dict(theta_start=0, theta_end=90, theta_samples=91, phi_start=0, phi_end=90, phi_samples=91, grid_size=1.6,
           n_x=267, n_y=267, max_bounces=10, freq=10e9)


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def sphere(ctx):
    v, f = M.icosphere(5)
    return v, f, ctx.mesh(v, f), O.Scene(v, f, 0), O.Scene(v, f, 1)


@pytest.fixture(scope="module")
def tri_duct(ctx):
    v, f = M.trihedral_duct_scene()
    return v, f, ctx.mesh(v, f), O.Scene(v, f, 0), O.Scene(v, f, 1)


def _step(msg):
    print(f"  [step] {msg}", flush=True)


def _db(z):
    return 10 * np.log10(np.maximum(4 * np.pi * np.abs(z) ** 2, 1e-10))


@pytest.mark.parametrize("case,angles", [("sphere", [0, 37, 200]), ("tri_duct", [0, 4140, 2790, 8280])])
def test_per_ray_fp32_bitexact_vs_oracle_twin(ctx, sphere, tri_duct, case, angles):
    v, f, mesh, s32, s64 = sphere if case == "sphere" else tri_duct
    kw = C1 if case == "sphere" else C2S
    p, op = S.SweepParams.make(**kw), O.Params(**kw)
    for a in angles:
        print(f"[per-ray fp32] {case} angle {a}", flush=True)
        g = ctx.trace_rays(mesh, p, a, S.SBR_FP32)
        o = s32.trace_angle(op, a, fused=True)
        assert np.array_equal(g["count"], o["count"])
        assert np.array_equal(g["ids"], o["ids"])
        assert np.array_equal(g["L"], o["L"])
        # FP32 vs FP64: only ties may differ
        o64 = s64.trace_angle(op, a)
        mm = PU.mismatches(g["ids"], g["count"], o64["ids"], o64["count"])
        # grazing aspects (a grid row lying in a plate's plane at theta = 90)
        # produce whole lines of ties; the fraction stays small
        assert len(mm) <= 1e-2 * len(g["count"])
        ties, real, rl = PU.classify(s64, op, a, mm, g["ids"], g["count"], limit=300)
        assert real == 0, f"non-tie mismatches at rays {rl[:10]}"


@pytest.mark.parametrize("case,angles", [("sphere", [0, 123]), ("tri_duct", [4140, 2790])])
def test_per_ray_fp64_bitexact_vs_oracle(ctx, sphere, tri_duct, case, angles):
    v, f, mesh, s32, s64 = sphere if case == "sphere" else tri_duct
    kw = C1 if case == "sphere" else C2S
    p, op = S.SweepParams.make(**kw), O.Params(**kw)
    for a in angles:
        print(f"[per-ray fp64] {case} angle {a}", flush=True)
        g = ctx.trace_rays(mesh, p, a, S.SBR_FP64)
        o = s64.trace_angle(op, a)
        assert np.array_equal(g["count"], o["count"])
        assert np.array_equal(g["ids"], o["ids"])
        assert np.array_equal(g["L"], o["L"])
        np.testing.assert_allclose(g["reim"], o["reim"], rtol=0, atol=1e-12 * np.abs(o["reim"]).max())


@pytest.mark.parametrize("case,angles", [("sphere", [0, 37]), ("tri_duct", [0, 4140, 2790, 8280, 7188, 4161])])
def test_per_ray_mixed_has_fp64_outcome(ctx, sphere, tri_duct, case, angles):
    """Mixed mode: every ray ends with the FP64 oracle's outcome -- queries
    decided in FP32 because none of their decisions was within its error
    bound, the others because they were answered by the FP64 query on the
    FP64 tree (flagged rays). The trihedral+duct angles include such rays."""
    v, f, mesh, s32, s64 = sphere if case == "sphere" else tri_duct
    kw = C1 if case == "sphere" else C2S
    p, op = S.SweepParams.make(**kw), O.Params(**kw)
    flagged = 0
    for a in angles:
        print(f"[per-ray mixed] {case} angle {a}", flush=True)
        g = ctx.trace_rays(mesh, p, a, S.SBR_MIXED)
        o = s64.trace_angle(op, a)
        assert np.array_equal(g["count"], o["count"])
        assert np.array_equal(g["ids"], o["ids"])
        # the FP64 ray state is the reference FP64 ray: path lengths are bit-exact
        assert np.array_equal(g["L"], o["L"])
        np.testing.assert_allclose(g["reim"], o["reim"], rtol=0, atol=1e-12 * np.abs(o["reim"]).max())
        # same path lengths and normals as the FP64 kernel; only the sin and
        # cos of the reduced phase are formed differently (the mixed mode's
        # series against the device library): a few units in the last place
        g64 = ctx.trace_rays(mesh, p, a, S.SBR_FP64)
        assert np.array_equal(g["L"], g64["L"])
        np.testing.assert_allclose(g["reim"], g64["reim"], rtol=0, atol=2e-15 * np.abs(g64["reim"]).max())
        rt = g["retraced"] == 1
        assert rt.mean() < 0.05
        flagged += int(rt.sum())
    print(f"    rays with an FP64-answered query: {flagged}", flush=True)
    if case == "tri_duct":
        assert flagged > 0, "the FP64 answer path was not exercised"


@pytest.mark.parametrize("case", ["sphere", "tri_duct"])
def test_per_ray_mixed_every_octant(ctx, sphere, tri_duct, case):
    """Launch directions in all eight octants and along the axes (exactly
    zero direction components, signed zeros included): the mixed query takes
    its per-octant box test whenever a wavefront's rays share their direction
    signs (every first query) and the general test otherwise; either way every
    ray ends with the FP64 oracle's outcome."""
    v, f, mesh, s32, s64 = sphere if case == "sphere" else tri_duct
    dirs = [(t, ph) for t in (30, 150) for ph in (45, 135, 225, 315)] + [(0, 0), (90, 0), (90, 90), (90, 180),
                                                                          (90, 270), (180, 0)]
    for th, ph in dirs:
        kw = dict(theta_start=th, theta_end=th, theta_samples=1, phi_start=ph, phi_end=ph, phi_samples=1,
                  grid_size=2.5 if case == "sphere" else 1.6, n_x=160, n_y=160, max_bounces=6, freq=10e9)
        p, op = S.SweepParams.make(**kw), O.Params(**kw)
        _step(f"{case} theta {th} phi {ph}")
        g = ctx.trace_rays(mesh, p, 0, S.SBR_MIXED)
        o = s64.trace_angle(op, 0)
        assert np.array_equal(g["count"], o["count"])
        assert np.array_equal(g["ids"], o["ids"])
        assert np.array_equal(g["L"], o["L"])
        assert np.count_nonzero(g["count"]) > 0


@pytest.mark.parametrize("scale", [0.37, 2.0])
def test_per_ray_mixed_scaled_model(ctx, scale):
    """A model scale that rounds the FP32 vertices (0.37) makes the mixed
    mode read the reference FP64 vertex records; a power of two (2.0) keeps
    every FP64 vertex equal to its FP32 one, so the FP32 record is widened.
    Both must give the FP64 oracle's rays, built at the same scale."""
    v, f = M.trihedral_duct_scene()
    mesh = ctx.mesh(v, f, scale=scale)
    s64 = O.Scene(v, f, 1, scale=scale)
    kw = dict(C2S, grid_size=1.6 * scale)
    p, op = S.SweepParams.make(**kw), O.Params(**kw)
    for a in (4140, 2790, 7188):
        print(f"[per-ray mixed] tri_duct scale {scale} angle {a}", flush=True)
        g = ctx.trace_rays(mesh, p, a, S.SBR_MIXED)
        o = s64.trace_angle(op, a)
        assert np.array_equal(g["count"], o["count"])
        assert np.array_equal(g["ids"], o["ids"])
        assert np.array_equal(g["L"], o["L"])
        np.testing.assert_allclose(g["reim"], o["reim"], rtol=0, atol=1e-12 * np.abs(o["reim"]).max())
    ctx.destroy_mesh(mesh)


def test_sphere_sigma_all_precisions(ctx, sphere):
    v, f, mesh, s32, s64 = sphere
    kw = dict(C1, phi_end=350, phi_samples=36)
    p, op = S.SweepParams.make(**kw), O.Params(**kw)
    _step("oracle fp64 sweep")
    ref, _ = s64.sweep(op)
    ref_db = _db(ref)
    _step("fp64 batched")
    z64, _, db64 = ctx.sweep(mesh, p, S.SBR_FP64, S.SBR_BATCHED)
    np.testing.assert_allclose(z64, ref, rtol=1e-9)
    for prec in (S.SBR_FP32, S.SBR_MIXED):
        _step(f"precision {prec} per-angle")
        za, _, dba = ctx.sweep(mesh, p, prec, S.SBR_PER_ANGLE)
        _step(f"precision {prec} batched")
        zb, _, dbb = ctx.sweep(mesh, p, prec, S.SBR_BATCHED)
        assert np.array_equal(za, zb), "per-angle and batched sums differ"
        assert np.max(np.abs(dbb - ref_db)) <= 0.1
    _step("repeatability")
    z2, _, _ = ctx.sweep(mesh, p, S.SBR_MIXED, S.SBR_BATCHED)
    z3, _, _ = ctx.sweep(mesh, p, S.SBR_MIXED, S.SBR_BATCHED)
    assert np.array_equal(z2, z3)


def test_tri_duct_sigma_mixed_vs_fp64(ctx, tri_duct):
    """The scene whose multi-bounce returns are nearly incoherent: FP32
    traversal alone drifts by ~0.5 dB from FP64 here; the mixed mode must not."""
    v, f, mesh, s32, s64 = tri_duct
    kw = dict(C2S, theta_samples=10, phi_samples=10)
    p, op = S.SweepParams.make(**kw), O.Params(**kw)
    _step("oracle fp64 sweep")
    ref, _ = s64.sweep(op)
    _step("mixed batched")
    zm, _, dbm = ctx.sweep(mesh, p, S.SBR_MIXED, S.SBR_BATCHED)
    np.testing.assert_allclose(zm, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
    assert np.max(np.abs(dbm - _db(ref))) <= 1e-6
    _step("mixed per-angle")
    zp, _, _ = ctx.sweep(mesh, p, S.SBR_MIXED, S.SBR_PER_ANGLE)
    assert np.array_equal(zm, zp)


@pytest.mark.parametrize("scene", ["tri_duct", "airframe"])
def test_mixed_undecided_queries(ctx, tri_duct, scene):
    """Queries the mixed query cannot vouch for (near-ties, hits on the edge
    of their box) are answered by the full FP64 query: both paths must occur
    here, and the sums must equal the FP64 path's (every ray has the FP64
    outcome; only the summation differs: FP32 two-sum pairs vs FP64)."""
    if scene == "tri_duct":
        v, f, mesh, s32, s64 = tri_duct
        p = S.SweepParams.make(**dict(C2S, theta_samples=6, phi_samples=6))
    else:
        from sagittasbr_amd import workloads as W

        wl = W.c3((3e9,))
        v, f = wl["meshes"][0]
        mesh = ctx.mesh(v, f)
        p = S.SweepParams.make(**dict(wl["sweeps"][0][1], theta_start=88, theta_end=92, theta_samples=3,
                                      phi_start=223, phi_end=243, phi_samples=3))
    _step(f"{scene} mixed")
    pl = ctx.plan(S.sweep_jobs(p, mesh, S.SBR_MIXED), S.SBR_MIXED, S.SBR_BATCHED)
    pl.run()
    pl.wait()
    zm, _, _ = pl.results()
    c = pl.count()
    pl.close()
    print(f"    rays {c['rays']}, with an undecided query {c['retraced']}, undecided queries {c['fp64_queries']}",
          flush=True)
    _step(f"{scene} fp64")
    z64, _, _ = ctx.sweep(mesh, p, S.SBR_FP64, S.SBR_BATCHED)
    assert 0 < c["retraced"] <= c["fp64_queries"] < 0.01 * c["queries"]
    np.testing.assert_allclose(zm, z64, rtol=1e-9, atol=1e-9 * np.abs(z64).max())
    if scene == "airframe":
        ctx.destroy_mesh(mesh)


@pytest.mark.parametrize("prec", [S.SBR_FP32, S.SBR_FP64, S.SBR_MIXED])
def test_job_sums_independent_of_neighbours(ctx, tri_duct, prec):
    """Each job's sum comes from its own pass partials in its own grid order:
    a job with no tiles (max_bounces 0) gives exactly zero, and the other
    jobs' sums are the same to the bit with or without it and in any order."""
    v, f, mesh, s32, s64 = tri_duct
    p = S.SweepParams.make(**dict(C2S, theta_samples=3, phi_samples=3))
    jobs = list(S.sweep_jobs(p, mesh, prec))
    _step(f"precision {prec} plain")
    z0, _, _ = ctx.sweep(mesh, p, prec, S.SBR_BATCHED)
    empty = S.Job(jobs[0].mesh, 0, jobs[0].n_x, jobs[0].n_y, jobs[0].freq, jobs[0].theta_deg, jobs[0].phi_deg,
                  jobs[0].grid_size, jobs[0].reflection_const)
    mixed = [jobs[4], empty] + jobs[::-1]
    _step(f"precision {prec} with an empty job, reversed")
    pl = ctx.plan((S.Job * len(mixed))(*mixed), prec, S.SBR_BATCHED)
    pl.run()
    pl.wait()
    z, _, _ = pl.results()
    pl.close()
    assert z[1] == 0
    assert z[0] == z0[4]
    assert np.array_equal(z[2:], z0[::-1])


def _tiny_meshes():
    tri = (np.array([[-0.5, -0.4, 0.0], [0.6, -0.3, 0.1], [0.0, 0.7, -0.05]], np.float32), np.array([[0, 1, 2]]))
    return [("one triangle", tri), ("plate 2 triangles", M.plate(1.0, 1.0, 1, 1)),
            ("plate 4 triangles", M.plate(1.0, 1.0, 1, 2)), ("box 12 triangles", M.box(0.5))]


@pytest.mark.parametrize("prec", [S.SBR_FP32, S.SBR_FP64, S.SBR_MIXED])
def test_tiny_meshes_root_leaf(ctx, prec):
    """Meshes of one to twelve triangles, whose tree root is a leaf or whose
    wide tree is a single record (nothing to stage in LDS): per ray against
    the oracle, batched with either cull method against per-angle, and the
    flat plate at normal incidence against sigma = 4 pi A^2 / lambda^2."""
    for name, (v, f) in _tiny_meshes():
        mesh = ctx.mesh(v, f)
        sc = O.Scene(v, f, 0) if prec == S.SBR_FP32 else O.Scene(v, f, 1)
        kw = dict(theta_start=0, theta_end=120, theta_samples=4, phi_start=10, phi_end=100, phi_samples=3,
                  grid_size=2.0, n_x=64, n_y=48, max_bounces=4, freq=5e9)
        p, op = S.SweepParams.make(**kw), O.Params(**kw)
        for a in (0, 4, 7):
            _step(f"precision {prec} {name} angle {a}: per ray")
            g = ctx.trace_rays(mesh, p, a, prec)
            o = sc.trace_angle(op, a, fused=prec == S.SBR_FP32)
            assert np.array_equal(g["count"], o["count"])
            assert np.array_equal(g["ids"], o["ids"])
            assert np.array_equal(g["L"], o["L"])
        jobs = S.sweep_jobs(p, mesh, prec)
        pa = ctx.plan(jobs, prec, S.SBR_PER_ANGLE)
        pa.run()
        pa.wait()
        z0, _, _ = pa.results()
        pa.close()
        assert np.count_nonzero(z0) > 0
        pl = ctx.plan(jobs, prec, S.SBR_BATCHED)
        for cull in (1, 2):
            _step(f"precision {prec} {name}: batched cull={cull}")
            pl.tune(cull=cull, top_levels=4)
            pl.run()
            pl.wait()
            z, _, _ = pl.results()
            assert np.array_equal(z, z0)
        pl.close()
        ctx.destroy_mesh(mesh)
    # all rays hit the plate with the same path length: |sum| = A / lambda
    v, f = M.plate(2.0, 2.0, 3, 5)
    mesh = ctx.mesh(v, f)
    g, n, freq = 1.2, 60, 3e9
    p = S.SweepParams.make(theta_start=0, theta_end=0, theta_samples=1, phi_start=0, phi_end=0, phi_samples=1,
                           grid_size=g, n_x=n, n_y=n, max_bounces=4, freq=freq)
    _, sig, _ = ctx.sweep(mesh, p, prec, S.SBR_BATCHED)
    lam = 299792458.0 / freq
    rel = 1e-12 if prec != S.SBR_FP32 else 1e-5
    assert sig[0] == pytest.approx(4 * np.pi * (g * g) ** 2 / lam ** 2, rel=rel)
    ctx.destroy_mesh(mesh)


@pytest.mark.parametrize("prec", [S.SBR_FP32, S.SBR_FP64, S.SBR_MIXED])
def test_non_square_ray_grid(ctx, tri_duct, prec):
    """n_x != n_y (the reference's two config keys, include/app/config.hpp:7-27):
    per ray against the oracle (the FP32 twin for FP32, the FP64 oracle
    otherwise), and the batched sums with either cull method equal to the
    per-angle sums, no bounce lost to the culling."""
    v, f, mesh, s32, s64 = tri_duct
    kw = dict(C2S, theta_samples=3, phi_samples=3, n_x=301, n_y=173)
    p, op = S.SweepParams.make(**kw), O.Params(**kw)
    for a in (0, 4, 8):
        _step(f"precision {prec} angle {a}: per ray, {p.n_x} x {p.n_y}")
        g = ctx.trace_rays(mesh, p, a, prec)
        o = s32.trace_angle(op, a, fused=True) if prec == S.SBR_FP32 else s64.trace_angle(op, a)
        assert np.array_equal(g["count"], o["count"])
        assert np.array_equal(g["ids"], o["ids"])
        assert np.array_equal(g["L"], o["L"])
        assert np.count_nonzero(g["count"]) > 1000
    jobs = S.sweep_jobs(p, mesh, prec)
    pa = ctx.plan(jobs, prec, S.SBR_PER_ANGLE)
    pa.run()
    pa.wait()
    z0, _, _ = pa.results()
    pa.close()
    pl = ctx.plan(jobs, prec, S.SBR_BATCHED)
    full = pl.count()
    for cull in (1, 2):
        for ts in (16, 64):
            _step(f"precision {prec} batched cull={cull} tile={ts}")
            pl.tune(cull=cull, tile_rays=ts)
            pl.run()
            pl.wait()
            z, _, _ = pl.results()
            assert np.array_equal(z, z0)
            tr = pl.count(traced=True)
            assert tr["ray_bounces"] == full["ray_bounces"] > 0
    pl.close()


@pytest.mark.parametrize("prec", [S.SBR_MIXED, S.SBR_FP32])
def test_plan_beyond_one_grid_row_of_jobs(ctx, sphere, prec):
    """A plan of more jobs than one launch grid dimension holds (65,535):
    the coverage and job-sum launches stride over the jobs. Jobs past the
    first 65,535 give the same sums as in a plan of their own, and the
    masks lose no bounce there."""
    v, f, mesh, s32, s64 = sphere
    p = S.SweepParams.make(theta_start=0, theta_end=180, theta_samples=190, phi_start=0, phi_end=359,
                           phi_samples=360, grid_size=2.5, n_x=16, n_y=16, max_bounces=2, freq=3e9)
    jobs = S.sweep_jobs(p, mesh, prec)
    n = len(jobs)
    assert n > 65535
    _step(f"precision {prec}: {n} jobs in one batched plan")
    pl = ctx.plan(jobs, prec, S.SBR_BATCHED)
    pl.run()
    pl.wait()
    z, _, _ = pl.results()
    full = pl.count()
    tr = pl.count(traced=True)
    pl.close()
    assert tr["ray_bounces"] == full["ray_bounces"] > 0
    for lo, hi in ((65530, n), (0, 2000)):
        _step(f"precision {prec}: jobs {lo}..{hi} on their own")
        sub = (S.Job * (hi - lo))(*[jobs[k] for k in range(lo, hi)])
        ps = ctx.plan(sub, prec, S.SBR_BATCHED)
        ps.run()
        ps.wait()
        zs, _, _ = ps.results()
        ps.close()
        assert np.array_equal(zs, z[lo:hi])
        assert np.all(np.abs(zs) > 0)


def test_tri_duct_sigma_fp64_and_methods(ctx, tri_duct):
    v, f, mesh, s32, s64 = tri_duct
    kw = dict(C2S, theta_samples=4, phi_samples=4, theta_end=90, phi_end=90)
    p, op = S.SweepParams.make(**kw), O.Params(**kw)
    _step("oracle fp64 sweep")
    ref, _ = s64.sweep(op)
    _step("fp64 per-angle")
    z64, _, _ = ctx.sweep(mesh, p, S.SBR_FP64, S.SBR_PER_ANGLE)
    np.testing.assert_allclose(z64, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
    _step("fp64 batched")
    z64b, _, _ = ctx.sweep(mesh, p, S.SBR_FP64, S.SBR_BATCHED)
    assert np.array_equal(z64, z64b)
    # FP32 kernel equals the oracle's FP32 twin summed in double, up to the
    # summation order and sin/cos of the two libraries
    _step("oracle fp32 sweep")
    r32, _ = s32.sweep(op, fused=True)
    _step("fp32 batched")
    z32, _, _ = ctx.sweep(mesh, p, S.SBR_FP32, S.SBR_BATCHED)
    np.testing.assert_allclose(z32, r32, rtol=2e-4, atol=2e-4 * np.abs(r32).max())


def test_duplicate_and_degenerate_triangles(ctx):
    """A sphere with a quarter of its faces repeated (every ray that hits one
    of them meets two triangles at exactly the same distance: the reference
    keeps the one its tree order tests last, rtKernels.cu:34-40) and three
    degenerate faces (a repeated vertex, twice, and three collinear vertices,
    rejected by the determinant test): per ray against the oracle in every
    precision. The mixed mode answers the exact ties with the FP64 query."""
    v, f = M.icosphere(3)
    v, f = np.asarray(v, np.float32), np.asarray(f, np.uint32)
    rng = np.random.default_rng(5)
    dup = f[rng.choice(len(f), len(f) // 4, replace=False)]
    nv = len(v)
    line = np.array([[0.2, 0.1, 1.5], [0.4, 0.3, 1.5], [0.6, 0.5, 1.5]], np.float32)
    degen = np.array([[0, 0, 1], [5, 7, 5], [nv, nv + 1, nv + 2]], np.uint32)
    V, F = np.concatenate([v, line]), np.concatenate([f, dup, degen]).astype(np.uint32)
    mesh = ctx.mesh(V, F)
    s32, s64 = O.Scene(V, F, 0), O.Scene(V, F, 1)
    kw = dict(theta_start=0, theta_end=90, theta_samples=2, phi_start=0, phi_end=90, phi_samples=2, grid_size=2.5,
              n_x=200, n_y=200, max_bounces=5, freq=3e9)
    p, op = S.SweepParams.make(**kw), O.Params(**kw)
    flagged = dup_wins = 0
    for a in range(4):
        o64 = s64.trace_angle(op, a)
        o32 = s32.trace_angle(op, a, fused=True)
        for prec in (S.SBR_FP32, S.SBR_FP64, S.SBR_MIXED):
            _step(f"angle {a} precision {prec}")
            g = ctx.trace_rays(mesh, p, a, prec)
            o = o32 if prec == S.SBR_FP32 else o64
            assert np.array_equal(g["count"], o["count"])
            assert np.array_equal(g["ids"], o["ids"])
            assert np.array_equal(g["L"], o["L"])
            if prec == S.SBR_MIXED:
                flagged += int(np.count_nonzero(g["retraced"]))
        # hits won by a repeated face (ids past the original faces)
        dup_wins += int(np.count_nonzero((o64["ids"] >= len(f)) & (o64["ids"] < len(f) + len(dup))))
    assert dup_wins > 0
    assert flagged > 0, "no exact tie reached the FP64 query"
    ctx.destroy_mesh(mesh)


def _corridor():
    """Two facing plates 0.2 m apart, 8 m deep below z = 0: rays entering at
    theta 28-35 degrees bounce between them 20 times and more."""
    va, fa = M._grid_quad([-0.1, -0.5, -8.0], [0, 1, 0], [0, 0, 8], 4, 32)
    vb, fb = M._grid_quad([0.1, -0.5, -8.0], [0, 0, 8], [0, 1, 0], 32, 4)
    return np.concatenate([va, vb]).astype(np.float32), np.concatenate([fa, fb + len(va)]).astype(np.uint32)


@pytest.mark.parametrize("prec", [S.SBR_FP32, S.SBR_FP64, S.SBR_MIXED])
def test_bounce_limits(ctx, prec):
    """max_bounces 1 (every hit ends its ray) and 20 (the reference's
    default, include/app/config.hpp:8-25), on two facing plates where rays
    reach 20 bounces: per ray against the oracle, and the batched sums
    against per-angle."""
    v, f = _corridor()
    mesh = ctx.mesh(v, f)
    s32, s64 = O.Scene(v, f, 0), O.Scene(v, f, 1)
    for mb in (1, 20):
        kw = dict(theta_start=28, theta_end=35, theta_samples=3, phi_start=0, phi_end=5, phi_samples=2,
                  grid_size=2.0, n_x=400, n_y=400, max_bounces=mb, freq=10e9)
        p, op = S.SweepParams.make(**kw), O.Params(**kw)
        top = 0
        for a in (0, 3, 5):
            _step(f"precision {prec} max_bounces {mb} angle {a}")
            g = ctx.trace_rays(mesh, p, a, prec)
            o = s32.trace_angle(op, a, fused=True) if prec == S.SBR_FP32 else s64.trace_angle(op, a)
            assert np.array_equal(g["count"], o["count"])
            assert np.array_equal(g["ids"], o["ids"])
            assert np.array_equal(g["L"], o["L"])
            assert g["count"].max() <= mb
            top += np.count_nonzero(g["count"] == mb)
        assert top > 0, "no ray reached the bounce limit"
        jobs = S.sweep_jobs(p, mesh, prec)
        sums = []
        for method in (S.SBR_PER_ANGLE, S.SBR_BATCHED):
            pl = ctx.plan(jobs, prec, method)
            pl.run()
            pl.wait()
            sums.append(pl.results()[0])
            pl.close()
        assert np.array_equal(sums[0], sums[1])
    ctx.destroy_mesh(mesh)


@pytest.mark.parametrize("refl", [0.5, 1.25])
def test_reflection_constant(ctx, tri_duct, refl):
    """The reflection constant R enters each ray's PO term as R^n over its n
    bounces (poKernels.cu:83-86): sigma of every precision against the
    oracle at the same R, and per ray, R scales the mixed mode's R = 1
    contribution by R^n (a property independent of the oracle)."""
    v, f, mesh, s32, s64 = tri_duct
    kw = dict(C2S, theta_samples=3, phi_samples=3, reflection_const=refl)
    p, op = S.SweepParams.make(**kw), O.Params(**kw)
    _step(f"R {refl}: oracle fp64 sweep")
    ref, _ = s64.sweep(op)
    for prec in (S.SBR_FP64, S.SBR_MIXED):
        _step(f"R {refl}: precision {prec} batched")
        z, _, _ = ctx.sweep(mesh, p, prec, S.SBR_BATCHED)
        np.testing.assert_allclose(z, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
    _step(f"R {refl}: oracle fp32 sweep and fp32 batched")
    r32, _ = s32.sweep(op, fused=True)
    z32, _, _ = ctx.sweep(mesh, p, S.SBR_FP32, S.SBR_BATCHED)
    np.testing.assert_allclose(z32, r32, rtol=2e-4, atol=2e-4 * np.abs(r32).max())
    p1 = S.SweepParams.make(**dict(kw, reflection_const=1.0))
    multi = 0
    for a in (0, 4, 8):
        _step(f"R {refl}: per ray, angle {a}")
        g = ctx.trace_rays(mesh, p, a, S.SBR_MIXED)
        g1 = ctx.trace_rays(mesh, p1, a, S.SBR_MIXED)
        assert np.array_equal(g["count"], g1["count"])
        scale = float(refl) ** g["count"].astype(np.float64)
        np.testing.assert_allclose(g["reim"], g1["reim"] * scale[:, None], rtol=1e-13,
                                   atol=1e-15 * np.abs(g1["reim"]).max())
        multi += np.count_nonzero(g["count"] > 1)
    assert multi > 0, "no multi-bounce ray exercised R^n"



def test_counters_match_oracle(ctx, tri_duct):
    v, f, mesh, s32, s64 = tri_duct
    kw = dict(C2S, theta_samples=1, theta_start=45, theta_end=45, phi_samples=1, phi_start=45, phi_end=45)
    p, op = S.SweepParams.make(**kw), O.Params(**kw)
    jobs = S.sweep_jobs(p, mesh, S.SBR_FP32)
    pl = ctx.plan(jobs, S.SBR_FP32, S.SBR_BATCHED)
    _step("fp32 batched count")
    c = pl.count()
    o = s32.trace_angle(op, 0, fused=True)
    culled = op.n_rays - c["rays"]
    assert c["tri_tests"] == o["tests"]
    assert c["node_visits"] + culled == o["visits"]
    assert c["ray_bounces"] == int(o["count"].sum())
    pl.close()


def test_graph_replay_matches_eager(ctx, sphere, tri_duct):
    """A captured sweep replayed several times, with other eager work and
    synchronisations between replays, gives the eager result every time."""
    v, f, mesh, s32, s64 = tri_duct
    kw = dict(C2S, theta_samples=6, phi_samples=6)
    p = S.SweepParams.make(**kw)
    for method in (S.SBR_BATCHED, S.SBR_PER_ANGLE):
        _step(f"mixed method {method} eager")
        jobs = S.sweep_jobs(p, mesh, S.SBR_MIXED)
        pl = ctx.plan(jobs, S.SBR_MIXED, method)
        pl.run()
        pl.wait()
        z0, _, _ = pl.results()
        pl.tune(use_graph=1)
        other = ctx.plan(S.sweep_jobs(S.SweepParams.make(**C1), sphere[2], S.SBR_FP32)[:3], S.SBR_FP32)
        for k in range(4):
            _step(f"graph replay {k}")
            pl.run()
            pl.wait()
            z, _, _ = pl.results()
            assert np.array_equal(z, z0)
            other.run()
            other.wait()
        other.close()
        pl.close()


def test_caller_stream(ctx, sphere):
    """Sweeps issued on a caller's stream (eager and hipGraph replay) give the
    sums of the context's own stream. The stream comes from the HIP runtime
    libsbr.so is linked against (torch's wheel carries a runtime of its own,
    whose streams are not valid there)."""
    import ctypes

    v, f, mesh, s32, s64 = sphere
    p = S.SweepParams.make(**dict(C1, phi_end=300, phi_samples=6))
    z0, _, _ = ctx.sweep(mesh, p, S.SBR_MIXED, S.SBR_BATCHED)
    hip = ctypes.CDLL("libamdhip64.so.7")
    st = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(st)) == 0
    ctx.set_stream(st.value)
    try:
        _step("sweep on the caller's stream")
        z1, _, _ = ctx.sweep(mesh, p, S.SBR_MIXED, S.SBR_BATCHED)
        pl = ctx.plan(S.sweep_jobs(p, mesh, S.SBR_MIXED), S.SBR_MIXED, S.SBR_BATCHED)
        pl.tune(use_graph=1)
        for _ in range(2):
            _step("graph replay on the caller's stream")
            pl.run()
            pl.wait()
            z2, _, _ = pl.results()
            assert np.array_equal(z2, z0)
        pl.close()
    finally:
        ctx.set_stream(None)
        hip.hipStreamDestroy(st)
    assert np.array_equal(z1, z0)
    z3, _, _ = ctx.sweep(mesh, p, S.SBR_MIXED, S.SBR_BATCHED)
    assert np.array_equal(z3, z0)


def test_graph_replay_after_mesh_table_grows():
    """A plan captured in a hipGraph holds the device address of the mesh
    table; meshes created later grow the table to a new address. The next
    replay must use the live table (captured again), not the freed one."""
    c = S.Context(0)
    try:
        v, f = M.icosphere(3)
        m0 = c.mesh(v, f)
        kw = dict(C1, phi_samples=8, phi_end=70, n_x=96, n_y=96)
        p = S.SweepParams.make(**kw)
        for prec in (S.SBR_FP32, S.SBR_MIXED):
            _step(f"graph capture, precision {prec}")
            pl = c.plan(S.sweep_jobs(p, m0, prec), prec, S.SBR_BATCHED)
            pl.tune(use_graph=1)
            pl.run()
            pl.wait()
            z0, _, _ = pl.results()
            for k in range(3):
                # new meshes and a plan on each grow both mesh tables
                vk, fk = M.icosphere(2 + k % 2)
                mk = c.mesh(vk * (1.0 + 0.1 * k), fk)
                other = c.plan(S.sweep_jobs(p, mk, S.SBR_MIXED)[:2], S.SBR_MIXED, S.SBR_BATCHED)
                other.run()
                other.wait()
                other.close()
                _step(f"replay after table growth {k}")
                pl.run()
                pl.wait()
                z, _, _ = pl.results()
                assert np.array_equal(z, z0)
            pl.close()
    finally:
        c.close()


@pytest.mark.parametrize("offset", [0.0, 4e3, 4e4])
def test_tile_cull_far_offset_mesh(ctx, offset):
    """Batched runs skip tiles and ray patches whose rays cannot reach the
    mesh (coverage masks). The FP32 rounding of the box projections grows
    with the mesh's distance from the origin in grid sizes; for a body far
    down the ray path the sums with the masks (a timed run, and a count of
    the traced rays) must equal those of a counting run, which traces every
    ray, and no bounce may be lost."""
    v, f = M.icosphere(3)
    grid = 2.5
    th, ph = np.radians(90.0), np.radians(20.0)
    d = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    vo = (v.astype(np.float64) - offset * grid * d).astype(np.float32)
    mesh = ctx.mesh(vo, f)
    # directions close enough that the far body stays inside the launch grid
    # (its projection moves by offset * grid * angle)
    dd = 2.0 if offset == 0 else float(np.degrees(0.3 / (offset * grid)))
    p = S.SweepParams.make(theta_start=90 - dd, theta_end=90 + dd, theta_samples=3, phi_start=20 - dd,
                           phi_end=20 + dd, phi_samples=3, grid_size=grid, n_x=256, n_y=256, max_bounces=3, freq=3e9)
    for prec in (S.SBR_FP32, S.SBR_MIXED):
        _step(f"offset {offset} grid sizes, precision {prec}")
        pl = ctx.plan(S.sweep_jobs(p, mesh, prec), prec, S.SBR_BATCHED)
        pl.tune(tile_rays=16)
        pl.run()
        pl.wait()
        z_cull, _, _ = pl.results()
        full = pl.count()
        z_all, _, _ = pl.results()
        tr = pl.count(traced=True)
        z_tr, _, _ = pl.results()
        pl.close()
        assert np.all(np.abs(z_all) > 0)
        assert np.array_equal(z_cull, z_all)
        assert np.array_equal(z_tr, z_all)
        assert tr["ray_bounces"] == full["ray_bounces"]
    ctx.destroy_mesh(mesh)


@pytest.mark.parametrize("prec", [S.SBR_FP32, S.SBR_MIXED, S.SBR_FP64])
def test_coverage_masks_keep_every_hitting_ray(ctx, tri_duct, prec):
    """Batched runs skip the 8x8 ray patches that their jobs' coverage masks
    rule out (or, tuned so, that the per-tile tests of 256 boxes rule out,
    as per-angle runs do). Counting the rays such a run traces must find
    every bounce that the count over all rays finds, over fewer rays, with
    the same sums, at every tile side; a timed run gives the same sums too."""
    v, f, mesh, s32, s64 = tri_duct
    p = S.SweepParams.make(**dict(C2S, theta_samples=4, phi_samples=4))
    pl = ctx.plan(S.sweep_jobs(p, mesh, prec), prec, S.SBR_BATCHED)
    full = pl.count()
    z_all, _, _ = pl.results()
    assert full["ray_bounces"] > 0
    for ts in (16, 32, 64):
        _step(f"precision {prec}, tile {ts}")
        pl.tune(tile_rays=ts)
        tr = pl.count(traced=True)
        z_tr, _, _ = pl.results()
        assert tr["ray_bounces"] == full["ray_bounces"]
        assert 0 < tr["rays"] < full["rays"]
        assert np.array_equal(z_tr, z_all)
        pl.run()
        pl.wait()
        z_run, _, _ = pl.results()
        assert np.array_equal(z_run, z_all)
        # the per-tile box tests instead of the masks: the same bounces and sums
        # (either may trace a few more rays than the other: the masks' ray
        # ranges carry one ray of margin on each side)
        pl.tune(tile_rays=ts, cull=2)
        tb = pl.count(traced=True)
        z_tb, _, _ = pl.results()
        assert tb["ray_bounces"] == full["ray_bounces"]
        assert 0 < tb["rays"] < full["rays"]
        assert np.array_equal(z_tb, z_all)
        pl.run()
        pl.wait()
        z_run, _, _ = pl.results()
        assert np.array_equal(z_run, z_all)
    pl.close()
    per = ctx.plan(S.sweep_jobs(p, mesh, prec), prec, S.SBR_PER_ANGLE)
    tp = per.count(traced=True)
    z_per, _, _ = per.results()
    per.close()
    assert tp["ray_bounces"] == full["ray_bounces"]
    assert tp["rays"] < full["rays"]
    assert np.array_equal(z_per, z_all)


@pytest.mark.parametrize("prec", [S.SBR_FP32, S.SBR_MIXED, S.SBR_FP64])
def test_tuning_variants(ctx, tri_duct, prec):
    """Partial sums are kept per 16x16 ray pass and reduced in a fixed order,
    so every launch shape, traversal schedule and both methods give the same
    sigma to the bit."""
    v, f, mesh, s32, s64 = tri_duct
    p = S.SweepParams.make(**dict(C2S, theta_samples=5, phi_samples=5))
    jobs = S.sweep_jobs(p, mesh, prec)
    pl = ctx.plan(jobs, prec, S.SBR_BATCHED)
    pl.run()
    pl.wait()
    z0, _, _ = pl.results()
    for pool in (1, 2):
        for ts in (16, 32, 64):
            for sd in (8, 16, 32):
                for wg in ((1, 0) if sd != 16 else (2,)):
                    _step(f"precision {prec} batched pool={pool} tile={ts} sd={sd} wg={wg}")
                    pl.tune(wg_per_cu=wg, lds_stack=sd, tile_rays=ts, ray_pool=pool)
                    pl.run()
                    pl.wait()
                    z, _, _ = pl.results()
                    assert np.array_equal(z, z0), (pool, ts, sd, wg)
    for trav in (1, 2):
        for pool in (1, 2):
            for ts in (16, 64):
                _step(f"precision {prec} batched traversal={trav} pool={pool} tile={ts}")
                pl.tune(traversal=trav, ray_pool=pool, tile_rays=ts)
                pl.run()
                pl.wait()
                z, _, _ = pl.results()
                assert np.array_equal(z, z0), (trav, pool, ts)
    # the mixed mode's top tree levels staged in LDS (ignored elsewhere)
    for top in (1, 3, 4):
        for trav in (1, 2):
            _step(f"precision {prec} batched top_levels={top} traversal={trav}")
            pl.tune(top_levels=top, traversal=trav, tile_rays=32)
            pl.run()
            pl.wait()
            z, _, _ = pl.results()
            assert np.array_equal(z, z0), (top, trav)
    # both cull methods
    for cull in (2, 1):
        for ts in (16, 64):
            _step(f"precision {prec} batched cull={cull} tile={ts}")
            pl.tune(cull=cull, tile_rays=ts, top_levels=4 if cull == 2 else 0)
            pl.run()
            pl.wait()
            z, _, _ = pl.results()
            assert np.array_equal(z, z0), (cull, ts)
    # the mixed mode's binary and eight-wide trees (ignored elsewhere); the
    # eight-wide tree's top 3 levels do not all fit the LDS copy
    for width in (2, 8, 4):
        for top in (0, 3):
            _step(f"precision {prec} batched node_width={width} top_levels={top}")
            pl.tune(node_width=width, top_levels=top, traversal=1 if top else 2, tile_rays=32)
            pl.run()
            pl.wait()
            z, _, _ = pl.results()
            assert np.array_equal(z, z0), (width, top)
    pl.close()
    pa = ctx.plan(jobs, prec, S.SBR_PER_ANGLE)
    for pool in (1, 2):
        for ts in (16, 64):
            trav = 1 if ts == 16 else 2
            _step(f"precision {prec} per-angle pool={pool} tile={ts} traversal={trav}")
            pa.tune(tile_rays=ts, ray_pool=pool, traversal=trav)
            pa.run()
            pa.wait()
            z, _, _ = pa.results()
            assert np.array_equal(z, z0), (pool, ts, trav)
    for width in (2, 8):
        _step(f"precision {prec} per-angle node_width={width}")
        pa.tune(node_width=width, top_levels=2)
        pa.run()
        pa.wait()
        z, _, _ = pa.results()
        assert np.array_equal(z, z0), width
    pa.close()


def test_autotune(ctx, tri_duct):
    v, f, mesh, s32, s64 = tri_duct
    p = S.SweepParams.make(**dict(C2S, theta_samples=4, phi_samples=6))
    jobs = S.sweep_jobs(p, mesh, S.SBR_MIXED)
    for method in (S.SBR_BATCHED, S.SBR_PER_ANGLE):
        pl = ctx.plan(jobs, S.SBR_MIXED, method)
        pl.run()
        pl.wait()
        z0, _, _ = pl.results()
        _step(f"autotune method {method}")
        best = pl.autotune(sample_jobs=8 if method == S.SBR_BATCHED else 0, repeats=1)
        print("   ", best, flush=True)
        # launch shapes, the other loop shape, then (mixed, no ray pool) the
        # top tree levels in LDS at 3 or 4 levels, also with an 8-entry stack,
        # then node widths 2 and 8, then (batched) the other cull method
        base = (36 + 1 + 1 if method == S.SBR_BATCHED else 24 + 1)
        assert best["candidates"] in {base + a + b for a in (0, 2, 4) for b in (0, 2)}
        assert best["tile_rays"] in (16, 32, 64) and best["lds_stack"] in (8, 16, 32) and best["best_ms"] > 0
        assert best["ray_pool"] in (1, 2) and best["traversal"] in (0, 1) and best["top_levels"] in (0, 3, 4)
        assert best["node_width"] in (0, 2, 8)
        assert best["cull"] in ((0, 2) if method == S.SBR_BATCHED else (0,))
        pl.run()
        pl.wait()
        z, _, _ = pl.results()
        assert np.array_equal(z, z0)
        pl.close()


@pytest.mark.parametrize("body", [2, 5])
def test_per_ray_mixed_cavity_bodies(ctx, body):
    """Mixed vs FP64 per ray on cylinders with a rectangular cavity (the
    random-bodies workload's multi-bounce kind), three angles each: hit
    faces, bounce counts and path lengths identical."""
    bodies = M.random_bodies(2, 6)
    b = bodies[body]
    assert b["kind"] == "cavity"
    mesh = ctx.mesh(b["verts"], b["faces"])
    r = float(np.max(np.linalg.norm(b["verts"], axis=1)))
    g = 2.02 * r
    n = int(np.ceil(g / (299792458.0 / b["freq"] / 10)))
    flagged = 0
    for th, ph in ((90.0, 0.0), (70.0, 33.0), (90.0, 90.0)):
        p = S.SweepParams.make(theta_start=th, theta_end=th, theta_samples=1, phi_start=ph, phi_end=ph, phi_samples=1,
                               grid_size=g, n_x=n, n_y=n, max_bounces=6, freq=b["freq"])
        _step(f"cavity body {body} ({len(b['faces'])} triangles, {n}x{n} rays) theta {th} phi {ph}")
        gm = ctx.trace_rays(mesh, p, 0, S.SBR_MIXED)
        g6 = ctx.trace_rays(mesh, p, 0, S.SBR_FP64)
        assert np.array_equal(gm["count"], g6["count"])
        assert np.array_equal(gm["ids"], g6["ids"])
        assert np.array_equal(gm["L"], g6["L"])
        flagged += int(gm["retraced"].sum())
    print(f"    rays with an FP64-answered query: {flagged}", flush=True)
    ctx.destroy_mesh(mesh)


@pytest.mark.parametrize("theta,phi", [(92.0, 223.0), (88.0, 230.0), (90.0, 242.0), (60.0, 217.0), (100.0, 240.0)])
def test_mixed_airframe_fin_angles(ctx, theta, phi):
    """Regression: on the airframe's fin two triangles are hit 8e-5 m apart
    and the farther one has a tiny determinant (distance error ~5e-3 m).
    The mixed mode must not decide such a query in FP32 (the competitor's
    own error bound counts) and must end with the FP64 outcome: per-ray
    bit-exact against the FP64 path (itself bit-exact with the FP64 oracle)."""
    from sagittasbr_amd import workloads as W

    wl = W.c3((3e9,))
    v, f = wl["meshes"][0]
    kw = dict(wl["sweeps"][0][1], theta_start=theta, theta_end=theta, theta_samples=1, phi_start=phi, phi_end=phi,
              phi_samples=1)
    p = S.SweepParams.make(**kw)
    mesh = ctx.mesh(v, f)
    _step(f"airframe 3 GHz theta {theta} phi {phi}: mixed")
    gm = ctx.trace_rays(mesh, p, 0, S.SBR_MIXED)
    _step("fp64")
    g6 = ctx.trace_rays(mesh, p, 0, S.SBR_FP64)
    ctx.destroy_mesh(mesh)
    assert np.array_equal(gm["count"], g6["count"])
    assert np.array_equal(gm["ids"], g6["ids"])
    assert np.array_equal(gm["L"], g6["L"])
    print(f"    rays {len(gm['count'])}, with an FP64-answered query {int(gm['retraced'].sum())}", flush=True)


def test_plan_outliving_context_fails_cleanly():
    c = S.Context(0)
    v, f = M.icosphere(2)
    mesh = c.mesh(v, f)
    p = S.SweepParams.make(**dict(C1, phi_samples=4, phi_end=3, n_x=32, n_y=32))
    pl = c.plan(S.sweep_jobs(p, mesh, S.SBR_FP32), S.SBR_FP32, S.SBR_BATCHED)
    pl.run()
    pl.wait()
    c.close()
    with pytest.raises(S.SbrError):
        pl.run()
    pl.close()


def test_bad_tuning_is_refused_and_leaves_the_plan_unchanged(ctx, tri_duct):
    v, f, mesh, s32, s64 = tri_duct
    p = S.SweepParams.make(**dict(C2S, theta_samples=2, phi_samples=3))
    jobs = S.sweep_jobs(p, mesh, S.SBR_MIXED)
    pl = ctx.plan(jobs, S.SBR_MIXED, S.SBR_BATCHED)
    pl.tune(tile_rays=32, top_levels=3)
    pl.run()
    pl.wait()
    z0, _, _ = pl.results()
    bad = [dict(node_width=3), dict(node_width=16), dict(top_levels=5), dict(top_levels=-1), dict(tile_rays=48),
           dict(traversal=3), dict(ray_pool=3), dict(lds_stack=-8), dict(wg_per_cu=-1)]
    for kw in bad:
        _step(f"refused tuning {kw}")
        with pytest.raises(S.SbrError, match="must be|negative"):
            pl.tune(**kw)
        pl.run()
        pl.wait()
        z, _, _ = pl.results()
        assert np.array_equal(z, z0), kw
    pl.close()
