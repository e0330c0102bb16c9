"""CPU tests of libsbr's host side: the C ABI loads and exports every symbol
of include/sbr.h, the host BVH equals the reference builder's output, the
launch frames equal the oracle's, and error paths return a status."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O
import sagittasbr_amd as S
from sagittasbr_amd import meshes as M

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "ref_bvh.npz"))


def test_library_exports_every_header_symbol():
    names = S.header_functions()
    assert len(names) >= 25
    L = S.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.sbr_abi_version() == 5


@pytest.mark.parametrize("name", ["ref_box", "ref_sphere", "ico3", "trihedral8", "random_soup"])
@pytest.mark.parametrize("prec,pi", [("f32", S.SBR_FP32), ("f64", S.SBR_FP64)])
def test_host_bvh_matches_reference_builder(name, prec, pi):
    v, f = GOLD[f"{name}_verts"], GOLD[f"{name}_faces"]
    h = S.host_bvh(v, f, precision=pi)
    assert np.array_equal(h["tris"], GOLD[f"{name}_{prec}_tris"])
    assert np.array_equal(h["bounds"], GOLD[f"{name}_{prec}_bounds"])
    assert np.array_equal(h["links"], GOLD[f"{name}_{prec}_links"])
    assert h["root"] == int(GOLD[f"{name}_{prec}_root"][0])


@pytest.mark.parametrize("parallel", [0, 1])
def test_host_bvh_matches_oracle_on_scene(parallel):
    v, f = M.trihedral_duct_scene()
    for pi in (S.SBR_FP32, S.SBR_FP64):
        h = S.host_bvh(v, f, precision=pi, options=S.BvhOptions.default(enable_parallel=parallel))
        sc = O.Scene(v, f, precision=pi)
        t, p, b, l = sc.export()
        assert np.array_equal(h["tris"], t) and np.array_equal(h["perm"], p)
        assert np.array_equal(h["bounds"], b) and np.array_equal(h["links"], l)


def test_host_bvh_options_median_and_leaf():
    v, f = M.icosphere(3)
    h = S.host_bvh(v, f, options=S.BvhOptions.default(algorithm=0, leaf_size=2))
    sc = O.Scene(v, f, precision=0, leaf_size=2, sah=False)
    t, p, b, l = sc.export()
    assert np.array_equal(h["links"], l) and np.array_equal(h["bounds"], b)
    assert l[:, 3].max() <= 2


def test_host_scale_applied_in_working_precision():
    v, f = M.icosphere(1)
    h = S.host_bvh(v, f, precision=S.SBR_FP32, scale=2.5)
    sc = O.Scene(v, f, precision=0, scale=2.5)
    assert np.array_equal(h["tris"], sc.export()[0])


def test_frames_match_oracle_bit_exact():
    kw = dict(theta_start=0, theta_end=180, theta_samples=19, phi_start=0, phi_end=359, phi_samples=37, grid_size=2.5,
              n_x=250, n_y=250)
    p, op = S.SweepParams.make(**kw), O.Params(**kw)
    for a in range(0, p.n_angles, 7):
        for pi in (S.SBR_FP32, S.SBR_FP64):
            fr, _ = S.angle_frame(p, a, pi)
            of, _ = O.frame(op, a, pi)
            assert np.array_equal(fr, of)


def test_sweep_jobs_enumeration_order():
    p = S.SweepParams.make(theta_start=10, theta_end=30, theta_samples=3, phi_start=0, phi_end=90, phi_samples=4)
    jobs = S.sweep_jobs(p, 0, S.SBR_FP32)
    th = [j.theta_deg for j in jobs]
    ph = [j.phi_deg for j in jobs]
    assert th == [10] * 4 + [20] * 4 + [30] * 4
    assert ph == [0, 30, 60, 90] * 3
    # single sample -> start value (sweep.cu:338-340)
    p1 = S.SweepParams.make(theta_start=45, theta_end=80, theta_samples=1, phi_samples=1, phi_start=7)
    j1 = S.sweep_jobs(p1, 0, S.SBR_FP64)
    assert j1[0].theta_deg == 45 and j1[0].phi_deg == 7


def test_error_paths_return_status():
    v, f = M.box()
    bad = f.copy()
    bad[0, 0] = 99
    with pytest.raises(S.SbrError, match="out-of-range"):
        S.host_bvh(v, bad)
    with pytest.raises(S.SbrError):
        S.host_bvh(v, f, options=S.BvhOptions.default(leaf_size=40))
    p = S.SweepParams.make(theta_samples=0)
    with pytest.raises(S.SbrError):
        S.sweep_jobs(p, 0)


def test_context_without_device_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(S.SbrError, match="no HIP device"):
        S.Context(0)
