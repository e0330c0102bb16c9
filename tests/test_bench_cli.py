"""CPU tests of bench.py's contract pieces that run without a GPU: the
one-GPU refusal, the argument checks, the tuning hand-over to the binding and
the dB error summary."""
import inspect
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402
from sagittasbr_amd import _native as N  # noqa: E402


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


@pytest.mark.parametrize("n", ["2", "8"])
def test_more_than_one_gpu_is_refused(n):
    r = _run("--gpus", n)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_profile_run_needs_a_tuning():
    r = _run("--profile-run")
    assert r.returncode != 0
    assert "--tuning" in r.stderr


def test_tuning_reaches_the_binding_in_its_order():
    """apply_tuning passes a tuning dict (as the autotuner returns it) to
    Plan.tune positionally: every field must land on its own parameter."""
    fields = [f for f, _ in N.Tuning._fields_]
    params = list(inspect.signature(N.Plan.tune).parameters)[1:]
    assert params == fields
    got = {}

    class FakePlan:
        def tune(self, *a):
            got.update(zip(fields, a))

    t = {f: 10 + k for k, f in enumerate(fields)}
    bench.apply_tuning(FakePlan(), dict(t, best_ms=1.0, candidates=3))
    assert got == t


def test_db_error_summary():
    ref = np.array([10.0, -5.0, -25.0, -50.0, 0.0])
    db = ref + np.array([0.01, -0.2, 0.05, 3.0, 0.0])
    s = bench.db_err(db, ref)
    assert s["angles_compared"] == 5
    assert s["max"] == pytest.approx(3.0)
    assert s["median"] == pytest.approx(0.05)
    # within 30 dB of the 10 dB peak: 10, -5, 0 (not -25 dB, which is 35 dB down)
    assert s["max_within_30dB_of_peak"] == pytest.approx(0.2)
    assert s["frac_angles_le_0.1dB"] == pytest.approx(3 / 5)
