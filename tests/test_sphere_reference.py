"""Sphere results the reference itself holds.

* assets/rcs_results.csv (reference): a PEC sphere of radius 1 m at 10 GHz,
  10000 x 10000 rays over 2.5 m, theta 90, 30 phi samples over 0..180 --
  sigma 3.1465 .. 3.1568 m^2 (mean 3.1514, 4.984 dBsm). Its mesh
  (sphere8.obj) is not in the repository; the test regenerates an
  icosphere of subdivision 8 (1,310,720 triangles, vertices on the sphere),
  so the two agree only to the mesh's own effect: within 0.05 dB per angle
  and on the mean.
* rcs_results.csv (reference, top level): at theta = 0 every phi launches
  the same rays (the frame does not depend on phi there), and the
  reference's 30 values spread by 2e-5 relative through its float atomics.
  Here sums are formed in a fixed order, so sigma at theta = 0 must be
  identical to the bit across phi, in every precision.
"""
import os

import numpy as np
import pytest

S = pytest.importorskip("sagittasbr_amd")
from sagittasbr_amd import meshes as M  # noqa: E402

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _step(msg):
    print(f"  [step] {msg}", flush=True)


def _reference_csv():
    rows = []
    with open(os.path.join(HERE, "golden", "ref_rcs_results_sphere_10GHz.csv")) as fh:
        for line in fh:
            if line.startswith("#") or line.startswith("theta"):
                continue
            rows.append([float(x) for x in line.strip().split(",")])
    return np.array(rows)


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


def test_sphere_10ghz_against_reference_results(ctx):
    ref = _reference_csv()
    assert ref.shape == (30, 4)
    _step("icosphere subdivision 8")
    v, f = M.icosphere(8)
    mesh = ctx.mesh(v, f)
    p = S.SweepParams.make(theta_start=90, theta_end=90, theta_samples=1, phi_start=0, phi_end=180, phi_samples=30,
                           grid_size=2.5, n_x=10000, n_y=10000, max_bounces=5, freq=1e10)
    for prec in (S.SBR_FP32, S.SBR_MIXED):
        _step(f"precision {prec}: 30 angles of 10000 x 10000 rays")
        _, sig, db = ctx.sweep(mesh, p, prec, S.SBR_BATCHED)
        d = db - ref[:, 3]
        mean_db = 10 * np.log10(sig.mean() / ref[:, 2].mean())
        print(f"    sigma {sig.min():.5f} .. {sig.max():.5f} m^2 (mean {sig.mean():.5f}); reference "
              f"{ref[:, 2].min():.5f} .. {ref[:, 2].max():.5f} (mean {ref[:, 2].mean():.5f}); per-angle dB "
              f"difference {d.min():+.4f} .. {d.max():+.4f}, mean {mean_db:+.4f} dB", flush=True)
        assert np.all(np.abs(d) <= 0.05)
        assert abs(mean_db) <= 0.05
    ctx.destroy_mesh(mesh)


def test_theta_zero_sigma_identical_across_phi(ctx):
    v, f = M.icosphere(5)
    mesh = ctx.mesh(v, f)
    p = S.SweepParams.make(theta_start=0, theta_end=0, theta_samples=1, phi_start=0, phi_end=348, phi_samples=30,
                           grid_size=2.5, n_x=250, n_y=250, max_bounces=5, freq=3e9)
    for prec in (S.SBR_FP32, S.SBR_FP64, S.SBR_MIXED):
        for method in (S.SBR_BATCHED, S.SBR_PER_ANGLE):
            _step(f"precision {prec} method {method}")
            z, sig, _ = ctx.sweep(mesh, p, prec, method)
            assert np.all(sig > 0)
            assert np.all(z == z[0]), np.unique(z)
    ctx.destroy_mesh(mesh)
