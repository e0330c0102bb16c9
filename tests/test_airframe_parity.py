"""The airframe (3,244,032 triangles; 1, 3 and 10 GHz at lambda/8, 8
bounces) against references independent of libsbr's own kernels.

  * per ray, one seeded angle per frequency over the whole launch grid
    (343,396 / 3,097,600 / 34,398,225 rays): libsbr's FP64 kernel and its
    mixed mode against the FP64 oracle -- hit face ids per bounce, bounce
    counts and path lengths bit-exact;
  * per ray, the same angles: libsbr's FP64 kernel against the reference's
    own USE_DOUBLE HIP build (rtKernels.cu built from its sources, with its
    own BVH builder) -- bounce counts and path lengths; a ray may differ only
    where it is an edge or grazing tie (its FP64 outcome changes under
    FP32-sized perturbations);
  * sigma of two further seeded angles per frequency from batched FP64 and
    mixed plans (their normal launches, with the per-tile emptiness test
    active, unlike the record and counting runs above) against the FP64
    oracle's sums (1e-9 relative).

These cover the four-wide collapse of the 30-level tree, the leaf encoding
at 3.2M triangles and the coverage masks of batched runs with checks that
do not go through another libsbr kernel. Counts of the traced rays against
counts over all rays show that the masks lose no bounce.
"""
import numpy as np
import pytest

import oracle_lib as O
import parity_util as PU

S = pytest.importorskip("sagittasbr_amd")
from sagittasbr_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu

FREQS = ("1 GHz", "3 GHz", "10 GHz")


def _step(msg):
    print(f"  [step] {msg}", flush=True)


@pytest.fixture(scope="module")
def air():
    wl = W.c3()
    v, f = wl["meshes"][0]
    ctx = S.Context(0)
    mesh = ctx.mesh(v, f)
    _step("FP64 oracle scene")
    s64 = O.Scene(v, f, 1)
    rng = np.random.default_rng(2026)
    # per sweep: one angle for the per-ray checks, two for the sigma checks
    picks = [rng.choice(kw["theta_samples"] * kw["phi_samples"], 3, replace=False) for _, kw in wl["sweeps"]]
    yield dict(wl=wl, v=v, f=f, ctx=ctx, mesh=mesh, s64=s64, picks=picks)
    ctx.close()


@pytest.mark.parametrize("k", [0, 1, 2], ids=FREQS)
def test_airframe_per_ray_vs_fp64_oracle(air, k):
    kw = air["wl"]["sweeps"][k][1]
    a = int(air["picks"][k][0])
    p, op = S.SweepParams.make(**kw), O.Params(**kw)
    ctx, mesh = air["ctx"], air["mesh"]
    _step(f"{FREQS[k]} angle {a}: FP64 oracle, {op.n_rays} rays")
    o = air["s64"].trace_angle(op, a)
    hit = int(np.count_nonzero(o["count"]))
    print(f"    rays hitting the body: {hit}, bounces: {int(o['count'].sum())}, max {int(o['count'].max())}",
          flush=True)
    assert hit > 1000
    for prec, name in ((S.SBR_FP64, "fp64"), (S.SBR_MIXED, "mixed")):
        _step(f"{FREQS[k]} angle {a}: libsbr {name} per ray")
        g = ctx.trace_rays(mesh, p, a, prec)
        assert np.array_equal(g["count"], o["count"])
        assert np.array_equal(g["ids"], o["ids"])
        assert np.array_equal(g["L"], o["L"])
        np.testing.assert_allclose(g["reim"], o["reim"], rtol=0, atol=1e-12 * np.abs(o["reim"]).max())
        if prec == S.SBR_MIXED:
            print(f"    mixed rays with an FP64-answered query: {int(np.count_nonzero(g['retraced']))}",
                  flush=True)
        del g


@pytest.mark.parametrize("k", [0, 1, 2], ids=FREQS)
def test_airframe_per_ray_vs_reference_fp64_build(air, k):
    if O.ref_gpu_lib("f64") is None:
        pytest.skip("reference GPU library not built (oracle/Makefile builds it where the reference tree exists)")
    kw = air["wl"]["sweeps"][k][1]
    a = int(air["picks"][k][0])
    p, op = S.SweepParams.make(**kw), O.Params(**kw)
    _step(f"{FREQS[k]} angle {a}: reference FP64 build, its own BVH")
    ref = O.RefGpu(air["v"], air["f"], "f64")
    try:
        rc, rd = ref.rays(op, a)
    finally:
        ref.close()
    _step(f"{FREQS[k]} angle {a}: libsbr FP64 per ray")
    g = air["ctx"].trace_rays(air["mesh"], p, a, S.SBR_FP64)
    same_cnt = rc == g["count"]
    same_len = np.abs(rd - g["L"]) <= 1e-9 * np.maximum(g["L"], 1.0)
    bad = np.nonzero(~(same_cnt & same_len))[0]
    exact = float(np.mean(rd[same_cnt] == g["L"][same_cnt]))
    print(f"    rays {len(rc)}, hitting {int(np.count_nonzero(rc))}, differing {len(bad)}, bit-identical path "
          f"lengths {exact:.6f}", flush=True)
    assert np.count_nonzero(rc) > 1000
    assert len(bad) <= 1e-4 * np.count_nonzero(rc) + 10
    _step(f"classifying {min(len(bad), 100)} differing rays")
    ties = sum(PU.is_tie(air["s64"], op, a, int(r)) for r in bad[:100])
    assert ties == min(len(bad), 100), f"non-tie differences at rays {bad[:10]}"


@pytest.mark.parametrize("k", [0, 1, 2], ids=FREQS)
def test_airframe_sigma_culled_vs_fp64_oracle(air, k):
    kw = air["wl"]["sweeps"][k][1]
    angles = np.sort(air["picks"][k][1:]).astype(np.int32)
    p, op = S.SweepParams.make(**kw), O.Params(**kw)
    ctx, mesh = air["ctx"], air["mesh"]
    _step(f"{FREQS[k]} angles {angles.tolist()}: FP64 oracle sums")
    ref, _ = air["s64"].sweep(op, angles)
    for prec, name in ((S.SBR_FP64, "fp64"), (S.SBR_MIXED, "mixed")):
        js = S.sweep_jobs(p, mesh, prec)
        sub = (S.Job * len(angles))(*[js[int(a)] for a in angles])
        for ts in (16, 64):
            _step(f"{FREQS[k]}: libsbr {name} batched, tile {ts}")
            pl = ctx.plan(sub, prec, S.SBR_BATCHED)
            try:
                pl.tune(tile_rays=ts)
                pl.run()
                pl.wait()
                z, _, _ = pl.results()
            finally:
                pl.close()
            np.testing.assert_allclose(z, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


@pytest.mark.parametrize("k", [0, 1, 2], ids=FREQS)
def test_airframe_coverage_masks_keep_every_hitting_ray(air, k):
    """Batched runs skip the 8x8 ray patches that the jobs' coverage masks
    (a 65,536-box cut of the airframe's tree) rule out: counting the rays
    such a run traces must find every bounce that the count over all rays
    of the culled rectangles finds, with the same sums."""
    kw = air["wl"]["sweeps"][k][1]
    angles = np.sort(air["picks"][k]).astype(np.int32)
    p = S.SweepParams.make(**kw)
    ctx, mesh = air["ctx"], air["mesh"]
    for prec, name in ((S.SBR_FP64, "fp64"), (S.SBR_MIXED, "mixed")):
        js = S.sweep_jobs(p, mesh, prec)
        sub = (S.Job * len(angles))(*[js[int(a)] for a in angles])
        pl = ctx.plan(sub, prec, S.SBR_BATCHED)
        try:
            _step(f"{FREQS[k]}: libsbr {name}, count over all rays")
            full = pl.count()
            z_all, _, _ = pl.results()
            for ts in (16, 64):
                _step(f"{FREQS[k]}: libsbr {name}, count over the traced rays, tile {ts}")
                pl.tune(tile_rays=ts)
                tr = pl.count(traced=True)
                z_tr, _, _ = pl.results()
                print(f"    traced rays {tr['rays']} of {full['rays']} ({tr['rays'] / full['rays']:.3f}), "
                      f"node visits {tr['node_visits'] / full['node_visits']:.3f}", flush=True)
                assert tr["ray_bounces"] == full["ray_bounces"] > 0
                assert tr["rays"] < full["rays"]
                assert np.array_equal(z_tr, z_all)
        finally:
            pl.close()
