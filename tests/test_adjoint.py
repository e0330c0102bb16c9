"""Frozen-path adjoint of sigma: d sigma / d vertex and d sigma / d f.

Reference for the gradients: central differences of the FP64 oracle along
the same frozen paths (hit faces fixed, bounce geometry recomputed with the
reference's FP64 operations). The oracle's frozen-path function is itself
pinned here on the CPU:
  * along the faces its own traversal found, it reproduces the traced sum;
  * flat plate at normal incidence: sigma = 4 pi A^2 / lambda^2 grows as f^2,
    so d sigma / d f = 2 sigma / f exactly;
  * moving the whole body along the incidence direction shifts every ray's
    phase by the same amount, so sigma is unchanged;
  * its ray-restricted central differences equal full recomputations.
"""
import numpy as np
import pytest

import oracle_lib as O
from sagittasbr_amd import meshes as M

C0 = 299792458.0


def _traced(v, f, dirs, grid, n, maxb, freq):
    """FP64 oracle traversal of each direction: ids, counts and sums."""
    sc = O.Scene(v, f, 1)
    ids, cnt, sums = [], [], []
    for th, ph in dirs:
        p = O.direction_params(th, ph, grid, freq, n, n, maxb)
        r = sc.trace_angle(p, 0)
        ids.append(np.ascontiguousarray(r["ids"], np.int32))
        cnt.append(r["count"].astype(np.int32))
        sums.append(r["sum"])
    return np.stack(ids), np.stack(cnt), np.array(sums)


def _verts64(v, scale=1.0):
    return np.asarray(v, np.float32).astype(np.float64) * scale


def _sigma(z):
    return 4 * np.pi * np.abs(z) ** 2


def _d0(th, ph):
    t, p = np.radians(th), np.radians(ph)
    return -np.array([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)])


def test_frozen_sum_reproduces_traced_paths():
    v, f = M.icosphere(3)
    dirs = [(90.0, 0.0), (37.0, 211.0), (120.0, 75.0)]
    ids, cnt, sums = _traced(v, f, dirs, 2.5, 64, 4, 3e9)
    pj, ij = O.job_arrays([d[0] for d in dirs], [d[1] for d in dirs], 2.5, 3e9, 64, 64, 4)
    for k in range(len(dirs)):
        z = O.frozen_sum(_verts64(v), f, pj[k], ij[k], ids[k], cnt[k])
        assert abs(z - sums[k]) <= 1e-12 * abs(sums[k])


def test_frozen_plate_frequency_closed_form():
    v, f = M.plate(1.0, 1.0, 4, 4)
    freq = 3e9
    ids, cnt, sums = _traced(v, f, [(0.0, 0.0)], 1.5, 64, 2, freq)
    assert cnt[0].max() == 1
    hf = 1e3
    s = []
    for fr in (freq - hf, freq, freq + hf):
        pj, ij = O.job_arrays([0.0], [0.0], 1.5, fr, 64, 64, 2)
        s.append(_sigma(O.frozen_sum(_verts64(v), f, pj[0], ij[0], ids[0], cnt[0])))
    fd = (s[2] - s[0]) / (2 * hf)
    assert abs(fd - 2 * s[1] / freq) <= 1e-7 * abs(2 * s[1] / freq)
    # and sigma itself is the plate closed form 4 pi A^2 / lambda^2 with A the
    # area of the rays that hit (N * dA)
    lam = C0 / freq
    area = cnt[0].sum() * (1.5 / 64) ** 2
    assert abs(s[1] - 4 * np.pi * area ** 2 / lam ** 2) <= 1e-9 * s[1]


def test_frozen_translation_along_incidence_is_invariant():
    v, f = M.icosphere(3)
    th, ph = 63.0, 141.0
    ids, cnt, sums = _traced(v, f, [(th, ph)], 2.5, 64, 4, 3e9)
    pj, ij = O.job_arrays([th], [ph], 2.5, 3e9, 64, 64, 4)
    d0 = _d0(th, ph)
    h = 1e-6
    sg = []
    for sgn in (-1, 0, 1):
        vv = _verts64(v) + sgn * h * d0
        sg.append(_sigma(O.frozen_sum(vv, f, pj[0], ij[0], ids[0], cnt[0])))
    along = abs(sg[2] - sg[0]) / (2 * h)
    # the same displacement across the incidence direction does change sigma
    u = np.cross(d0, [0.0, 0.0, 1.0])
    u /= np.linalg.norm(u)
    sx = [_sigma(O.frozen_sum(_verts64(v) + sgn * h * u, f, pj[0], ij[0], ids[0], cnt[0])) for sgn in (-1, 1)]
    across = abs(sx[1] - sx[0]) / (2 * h)
    assert across > 1e-3 * sg[1]
    assert along <= 1e-6 * across


def test_frozen_fd_restriction_equals_full_recompute():
    v, f = M.icosphere(2)
    dirs = [(90.0, 10.0), (45.0, 300.0)]
    ids, cnt, _ = _traced(v, f, dirs, 2.5, 48, 3, 3e9)
    pj, ij = O.job_arrays([d[0] for d in dirs], [d[1] for d in dirs], 2.5, 3e9, 48, 48, 3)
    w = np.array([0.7, -1.3])
    touched = np.unique(f[np.unique(ids[ids >= 0])].reshape(-1))
    coords = np.array([3 * touched[0], 3 * touched[len(touched) // 2] + 1, 3 * touched[-1] + 2], np.int32)
    h = 1e-6
    g, _ = O.frozen_fd(_verts64(v), f, pj, ij, ids, cnt, w, coords, h)
    for c, gc in zip(coords, g):
        tot = 0.0
        for sgn in (-1, 1):
            vv = _verts64(v).copy()
            vv[c // 3, c % 3] += sgn * h
            tot += sgn * sum(w[k] * _sigma(O.frozen_sum(vv, f, pj[k], ij[k], ids[k], cnt[k])) for k in range(2))
        full = tot / (2 * h)
        assert abs(full - gc) <= 1e-7 * max(abs(full), 1e-12) + 1e-9


def test_frozen_fd_step_convergence_multibounce():
    """On grazing multi-bounce paths the plain central difference at 1e-6 m
    carries a visible truncation error; the Richardson combination at 1e-7 m
    agrees with a plain 1e-8 m step (and with 2e-8 m Richardson)."""
    v, f = M.trihedral_duct_scene()
    dirs = [(55.0, 45.0)]
    ids, cnt, _ = _traced(v, f, dirs, 1.6, 96, 10, 10e9)
    assert cnt.max() >= 3
    pj, ij = O.job_arrays([55.0], [45.0], 1.6, 10e9, 96, 96, 10)
    multi = np.where(cnt[0] >= 3)[0]
    faces = np.unique(ids[0][multi][:, :3].reshape(-1))
    verts = np.unique(f[faces].reshape(-1))[:12]
    coords = np.concatenate([3 * verts + a for a in range(3)]).astype(np.int32)
    w = np.ones(1)
    g7r, _ = O.frozen_fd(_verts64(v), f, pj, ij, ids, cnt, w, coords, 1e-7, True)
    g8, _ = O.frozen_fd(_verts64(v), f, pj, ij, ids, cnt, w, coords, 1e-8)
    g8r, _ = O.frozen_fd(_verts64(v), f, pj, ij, ids, cnt, w, coords, 2e-8, True)
    scale = np.abs(g8r).max()
    assert scale > 0
    assert np.max(np.abs(g7r - g8r)) <= 1e-6 * scale
    assert np.max(np.abs(g8 - g8r)) <= 1e-5 * scale


# ----------------------------------------------------------------- GPU
S = pytest.importorskip("sagittasbr_amd")


def _rel_err(a, b, floor):
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def _gpu_vs_fd(ctx, v, f, dirs, grid, n, maxb, freq, n_coords, h, seed, w=None, richardson=False):
    mesh = ctx.mesh(v, f)
    th = np.array([d[0] for d in dirs])
    ph = np.array([d[1] for d in dirs])
    if w is None:
        w = np.random.default_rng(seed).normal(size=len(dirs))
    jobs = S.direction_jobs(mesh, th, ph, freq, grid, n, n, maxb)
    sig, dxyz, dsdf = ctx.adjoint(mesh, jobs, w)
    ids, cnt, sums = _traced(v, f, dirs, grid, n, maxb, freq)
    # the mixed-precision forward ends every ray with the FP64 outcome
    np.testing.assert_allclose(sig, _sigma(sums), rtol=1e-9)
    pj, ij = O.job_arrays(th, ph, grid, freq, n, n, maxb)
    touched = np.unique(f[np.unique(ids[ids >= 0])].reshape(-1))
    rng = np.random.default_rng(seed)
    coords = (3 * rng.choice(touched, n_coords, replace=len(touched) < n_coords) + rng.integers(0, 3, n_coords))
    coords = coords.astype(np.int32)
    g_fd, _ = O.frozen_fd(_verts64(v), f, pj, ij, ids, cnt, w, coords, h, richardson)
    g_ad = dxyz.reshape(-1)[coords]
    ctx.destroy_mesh(mesh)
    return g_ad, g_fd, sig, dsdf, pj, ij, ids, cnt


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


@pytest.mark.gpu
def test_adjoint_vertices_match_fd_sphere(ctx):
    v, f = M.icosphere(4)
    dirs = [(90.0, 0.0), (30.0, 45.0), (150.0, 200.0), (75.0, 310.0), (10.0, 10.0), (120.0, 100.0)]
    print("[adjoint] sphere subdiv 4, 6 directions", flush=True)
    g_ad, g_fd, *_ = _gpu_vs_fd(ctx, v, f, dirs, 2.5, 96, 4, 3e9, 64, 1e-5, seed=5)
    scale = np.abs(g_fd).max()
    err = _rel_err(g_ad, g_fd, 1e-3 * scale)
    assert err.max() <= 1e-3, (err.max(), g_ad[:5], g_fd[:5])


@pytest.mark.gpu
def test_adjoint_mixed_job_shapes_add_up(ctx):
    """One call over jobs of different grids, bounce limits and frequencies
    (their records lie one after another) equals the weighted sum of
    one-job calls: the adjoint is linear in the weights, and each job's sigma
    and d sigma / d f are its own."""
    v, f = M.trihedral_duct_scene()
    mesh = ctx.mesh(v, f)
    spec = [(55.0, 45.0, 10e9, 96, 10), (40.0, 30.0, 3e9, 64, 3), (70.0, 60.0, 5e9, 80, 0), (20.0, 80.0, 8e9, 128, 6)]
    jobs = [S.direction_jobs(mesh, [th], [ph], fr, 1.6, n, n, mb)[0] for th, ph, fr, n, mb in spec]
    w = np.array([0.5, -1.25, 2.0, 0.75])
    print("[adjoint] trihedral+duct, 4 jobs of different shapes in one call", flush=True)
    sg, gx, df = ctx.adjoint(mesh, jobs, w)
    gsum = np.zeros_like(gx)
    for k, j in enumerate(jobs):
        s1, g1, d1 = ctx.adjoint(mesh, [j], w[k:k + 1])
        assert s1[0] == sg[k]  # fixed-order sums
        # d sigma / d f and the vertex gradients are FP64 atomic sums
        assert abs(d1[0] - df[k]) <= 1e-12 * abs(df[k])
        gsum += g1
    assert sg[2] == 0.0 and df[2] == 0.0  # max_bounces 0: no hit, no gradient
    np.testing.assert_allclose(gx, gsum, rtol=0, atol=1e-12 * np.abs(gsum).max())
    ctx.destroy_mesh(mesh)


@pytest.mark.gpu
def test_adjoint_multibounce_matches_fd(ctx):
    """Trihedral + duct: up to 10 bounces, every bounce's geometry in the
    chain. Grazing multi-bounce paths make the path length move ~200x faster
    than a vertex, so the finite differences use h = 1e-7 m with Richardson
    extrapolation (a plain 1e-6 m step is off by ~1e-3 from truncation)."""
    v, f = M.trihedral_duct_scene()
    dirs = [(55.0, 45.0), (40.0, 30.0), (70.0, 60.0), (20.0, 80.0)]
    print("[adjoint] trihedral+duct, 4 directions", flush=True)
    g_ad, g_fd, _, _, pj, ij, ids, cnt = _gpu_vs_fd(ctx, v, f, dirs, 1.6, 128, 10, 10e9, 64, 1e-7, seed=7,
                                                    richardson=True)
    assert cnt.max() >= 3
    scale = np.abs(g_fd).max()
    err = _rel_err(g_ad, g_fd, 1e-3 * scale)
    assert err.max() <= 1e-3, (err.max(), g_ad[:5], g_fd[:5])


@pytest.mark.gpu
def test_adjoint_frequency(ctx):
    v, f = M.icosphere(4)
    dirs = [(90.0, 0.0), (30.0, 45.0), (150.0, 200.0)]
    print("[adjoint] frequency derivative", flush=True)
    _, _, sig, dsdf, pj, ij, ids, cnt = _gpu_vs_fd(ctx, v, f, dirs, 2.5, 96, 4, 3e9, 4, 1e-5, seed=9)
    hf = 1e3
    for k in range(len(dirs)):
        s = []
        for sgn in (-1, 1):
            pk = pj[k].copy()
            pk[3] += sgn * hf
            s.append(_sigma(O.frozen_sum(_verts64(v), f, pk, ij[k], ids[k], cnt[k])))
        fd = (s[1] - s[0]) / (2 * hf)
        assert abs(dsdf[k] - fd) <= 1e-6 * abs(fd) + 1e-12 * sig[k] / 3e9
    # flat plate at normal incidence: d sigma / d f = 2 sigma / f
    vp, fp = M.plate(1.0, 1.0, 4, 4)
    mesh = ctx.mesh(vp, fp)
    jobs = S.direction_jobs(mesh, [0.0], [0.0], 3e9, 1.5, 64, 64, 2)
    sg, _, df = ctx.adjoint(mesh, jobs, [1.0])
    assert abs(df[0] - 2 * sg[0] / 3e9) <= 1e-9 * abs(df[0])
    ctx.destroy_mesh(mesh)


@pytest.mark.gpu
def test_adjoint_translation_invariance(ctx):
    """sum_v d sigma / d v . d0 = 0: translating the body along the incidence
    direction shifts every phase equally."""
    v, f = M.icosphere(4)
    mesh = ctx.mesh(v, f)
    for th, ph in [(63.0, 141.0), (90.0, 0.0), (20.0, 250.0)]:
        jobs = S.direction_jobs(mesh, [th], [ph], 3e9, 2.5, 128, 128, 4)
        _, g, _ = ctx.adjoint(mesh, jobs, [1.0])
        d0 = _d0(th, ph)
        proj = g @ d0
        assert abs(proj.sum()) <= 1e-9 * np.abs(proj).sum(), (th, ph, proj.sum(), np.abs(proj).sum())
    ctx.destroy_mesh(mesh)


@pytest.mark.gpu
def test_adjoint_c5_sh_sphere(ctx):
    """The differentiation workload: SH-perturbed sphere (icosphere 7,
    327,680 triangles, 2 cm RMS), 3 GHz, lambda/10 (2.5 m, 250 x 250), max 4
    bounces, 64 directions uniform on the sphere; 256 random vertex
    coordinates against FP64 central differences with h = 1e-5 m (Richardson
    combination of the 1e-5 and 5e-6 steps)."""
    v, f, _ = M.sh_perturbed_sphere(7, seed=3)
    rng = np.random.default_rng(3)
    u, w_ = rng.random(64), rng.random(64)
    dirs = list(zip(np.degrees(np.arccos(1 - 2 * u)), 360.0 * w_))
    print("[adjoint] C5 SH sphere, 64 directions, 256 coordinates", flush=True)
    g_ad, g_fd, sig, *_ = _gpu_vs_fd(ctx, v, f, dirs, 2.5, 250, 4, 3e9, 256, 1e-5, seed=3,
                                    w=np.ones(64) / 64, richardson=True)
    scale = np.abs(g_fd).max()
    err = _rel_err(g_ad, g_fd, 1e-3 * scale)
    print(f"    max rel err {err.max():.3e}, median {np.median(err):.3e}", flush=True)
    assert err.max() <= 1e-3


@pytest.mark.gpu
def test_autograd_function(ctx):
    import torch

    from sagittasbr_amd.adjoint import monostatic_rcs

    v, f = M.icosphere(3)
    vt = torch.tensor(v.astype(np.float64), requires_grad=True)
    ft = torch.tensor(3e9, dtype=torch.float64, requires_grad=True)
    th, ph = np.array([90.0, 45.0]), np.array([0.0, 120.0])
    sigma = monostatic_rcs(ctx, vt, f, th, ph, ft, grid_size=2.5, n_x=96, n_y=96, max_bounces=4)
    w = torch.tensor([0.3, -1.1], dtype=torch.float64)
    (w * sigma).sum().backward()
    mesh = ctx.mesh(v, f)
    sg, gx, df = ctx.adjoint(mesh, S.direction_jobs(mesh, th, ph, 3e9, 2.5, 96, 96, 4), w.numpy())
    ctx.destroy_mesh(mesh)
    np.testing.assert_allclose(sigma.detach().numpy(), sg, rtol=1e-12)
    np.testing.assert_allclose(vt.grad.numpy(), gx, rtol=1e-9, atol=1e-12 * np.abs(gx).max())
    assert abs(ft.grad.item() - float(w.numpy() @ df)) <= 1e-9 * abs(float(w.numpy() @ df))


@pytest.mark.gpu
def test_autograd_mesh_cache_and_single_adjoint(ctx):
    """Forward runs one sweep on a cached mesh; backward runs one adjoint.
    Re-evaluating a geometry reuses its mesh, at most two stay alive, and a
    backward pass whose mesh was evicted builds it again with the same
    gradient."""
    import torch

    from sagittasbr_amd import adjoint as A

    v, f = M.icosphere(3)
    th, ph = np.array([90.0, 30.0]), np.array([10.0, 200.0])
    kw = dict(grid_size=2.5, n_x=80, n_y=80, max_bounces=4)
    eng = S.Context(0)
    try:
        vt = torch.tensor(v.astype(np.float64), requires_grad=True)
        s1 = A.monostatic_rcs(eng, vt, f, th, ph, 3e9, **kw)
        s2 = A.monostatic_rcs(eng, vt, f, th, ph, 3e9, **kw)
        assert len(eng._adjoint_meshes) == 1
        assert torch.equal(s1, s2)
        s1.sum().backward()
        g_first = vt.grad.clone()
        # two other geometries push the first mesh out of the cache
        for k in (1, 2):
            vk = torch.tensor((v * (1 + 0.01 * k)).astype(np.float64), requires_grad=True)
            A.monostatic_rcs(eng, vk, f, th, ph, 3e9, **kw)
        assert len(eng._adjoint_meshes) == 2
        vt.grad = None
        s2.sum().backward()  # its mesh was evicted: rebuilt for the backward pass
        # (FP64 atomics: equal to rounding)
        np.testing.assert_allclose(vt.grad.numpy(), g_first.numpy(), rtol=1e-12,
                                   atol=1e-13 * np.abs(g_first.numpy()).max())
        assert len(eng._adjoint_meshes) == 2
    finally:
        eng.close()
