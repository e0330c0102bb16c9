#!/usr/bin/env python3
"""Benchmark: the batched multi-bounce sweep on one MI355X.

Workload (one step): trihedral corner reflector + open duct (34,816
triangles), f = 10 GHz, lambda/10 ray spacing (1.6 m launch grid, 535 x 535
rays), max 10 bounces, theta 0..90 x phi 0..90 in 1 degree steps (8,281
angles), all angles in one persistent launch, mixed precision (FP32
traversal with error bounds, FP64 ray state and phase, queries that FP32
cannot decide answered on the FP64 tree). Mesh, BVH and the job list are resident in HBM before the
timed region; the launch shape is autotuned before it.

Prints one JSON line: angles/s, ray-bounces/s, the plain per-angle FP32 and
the batched FP32 / FP64 rates on the same GPU, accuracy vs FP64, the
roofline (algorithmic bytes from the FP64 traversal counters; HBM traffic
and VALU activity from rocprofv3 counter passes of the same kernel run in
child processes) and a CPU baseline (the oracle on the host cores).
"""
import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NODE_BYTES = 32  # 6 floats of bounds + 2 ints, per node visit of the reference traversal
TRI_BYTES = 36   # 9 floats, per Moller-Trumbore test
OUT_BYTES = 16   # one complex per angle
COPY_BYTES = 1 << 30  # calibration copy: 1 GiB read + 1 GiB written per launch
DOMINANT = "trace_kernel"

WORKLOAD = dict(theta_start=0.0, theta_end=90.0, theta_samples=91, phi_start=0.0, phi_end=90.0, phi_samples=91,
                grid_size=1.6, n_x=535, n_y=535, max_bounces=10, freq=10e9, reflection_const=1.0)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="mixed", choices=["fp32", "fp64", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU time of the baseline sample")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-pmc", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--tuning", default="", help="JSON tuning to apply instead of autotuning")
    ap.add_argument("--profile-run", action="store_true",
                    help="only the timed batched sweeps (for counter collection under rocprofv3)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB device copy (rocprof FETCH/WRITE_SIZE calibration)")
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5"],
                    help="c2 (default) is the metric's configuration; c1, c3, c4 are the other configurations")
    ap.add_argument("--stride", type=int, default=0,
                    help="c1/c3/c4: every stride-th angle for the FP64 and per-angle comparisons (0 = default)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def copy_calibration():
    import torch

    n = COPY_BYTES // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": COPY_BYTES, "copy_bytes_written": COPY_BYTES, "launches": 3}))


def cpu_baseline(verts, faces, kw, target_s):
    """The oracle (a C++/OpenMP restatement of the reference path; the
    reference has no CPU ray path) timed on the host cores over a bounded
    sample of the same sweep's angles."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    p = O.Params(**{k: v for k, v in kw.items()})
    sc = O.Scene(verts, faces, precision=0)
    n = p.n_angles
    # evenly spread probe angles; grow the sample until ~target_s
    order = np.random.default_rng(7).permutation(n).astype(np.int32)
    t0 = time.perf_counter()
    sc.sweep(p, order[:2])
    per = (time.perf_counter() - t0) / 2
    m = int(max(2, min(n, target_s / max(per, 1e-6))))
    sample = order[:m]
    t0 = time.perf_counter()
    sc.sweep(p, sample)
    dt = time.perf_counter() - t0
    return {"value": m / dt, "unit": "angles/s", "cores": O.lib().orc_num_threads(), "kind": "port",
            "sample": f"{m} of {n} angles (seeded random), FP32 oracle (reference arithmetic), {dt:.1f} s"}


def _pmc_pass(counters, extra_args, tmp, name):
    """One rocprofv3 counter pass over a child bench process; returns
    {kernel_name: [values per dispatch]} summed over counter instances."""
    out = os.path.join(tmp, name)
    cmd = ["rocprofv3", "--pmc", *counters, "--output-format", "csv", "-d", out, "-o", name, "--",
           sys.executable, os.path.join(ROOT, "bench.py"), *extra_args]
    env = dict(os.environ, TMPDIR="/tmp")
    r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=180)
    files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
    if r.returncode != 0 and not files:
        raise RuntimeError(f"rocprofv3 {name} exited {r.returncode}: {r.stdout[-400:]!r}")
    res = {}
    for f in files:
        for row in csv.DictReader(open(f)):
            key = (row["Kernel_Name"], row["Counter_Name"])
            res.setdefault(key, {}).setdefault(row["Dispatch_Id"], 0.0)
            res[key][row["Dispatch_Id"]] += float(row["Counter_Value"])
    return res


def _per_dispatch(res, kernel_sub, counter):
    vals = [v for (k, c), d in res.items() if kernel_sub in k and c == counter for v in d.values()]
    if not vals:
        raise RuntimeError(f"no {counter} rows for {kernel_sub}")
    return float(np.mean(vals))


def pmc_measure(tuning):
    """HBM bytes per launch of the dominant kernel (FETCH_SIZE and WRITE_SIZE,
    each scaled by the same counter's reading for a 1 GiB device copy) and
    its VALU busy / utilisation / occupancy."""
    if shutil.which("rocprofv3") is None:
        return {"error": "rocprofv3 not found"}
    tmp = tempfile.mkdtemp(prefix="sbr_pmc_", dir="/tmp")
    prof = ["--profile-run", "--steps", "1", "--warmup", "0", "--tuning", json.dumps(tuning)]
    try:
        f_b = _per_dispatch(_pmc_pass(["FETCH_SIZE"], prof, tmp, "fetch"), DOMINANT, "FETCH_SIZE")
        w_b = _per_dispatch(_pmc_pass(["WRITE_SIZE"], prof, tmp, "write"), DOMINANT, "WRITE_SIZE")
        f_c = _per_dispatch(_pmc_pass(["FETCH_SIZE"], ["--copy-calibration"], tmp, "cfetch"), "copyBuffer",
                            "FETCH_SIZE")
        w_c = _per_dispatch(_pmc_pass(["WRITE_SIZE"], ["--copy-calibration"], tmp, "cwrite"), "copyBuffer",
                            "WRITE_SIZE")
        v = _pmc_pass(["VALUBusy", "VALUUtilization"], prof, tmp, "valu")
        o = _pmc_pass(["MeanOccupancyPerCU"], prof, tmp, "occ")
        # counters are in KiB per dispatch; scale each by copy bytes / copy reading
        fs = COPY_BYTES / (f_c * 1024.0)
        ws = COPY_BYTES / (w_c * 1024.0)
        read = f_b * 1024.0 * fs
        write = w_b * 1024.0 * ws
        return {"read_bytes": read, "write_bytes": write, "fetch_scale": fs, "write_scale": ws,
                "valu_busy_pct": _per_dispatch(v, DOMINANT, "VALUBusy"),
                "valu_utilization_pct": _per_dispatch(v, DOMINANT, "VALUUtilization"),
                "mean_occupancy_waves_per_cu": _per_dispatch(o, DOMINANT, "MeanOccupancyPerCU")}
    except Exception as e:  # counters are reported beside the measurement, never instead of it
        return {"error": str(e)[:300]}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def timed(plan, steps, warmup):
    import torch

    for _ in range(warmup):
        plan.run()
        plan.wait()
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.run()
        plan.wait()
        kernel_ms.append(plan.last_ms())
    torch.cuda.synchronize()
    return 1000.0 * (time.perf_counter() - t0) / steps, kernel_ms


def adjoint_workload(args):
    """c5: batched mixed forward of the 64 directions, then the frozen-path
    adjoint (recording forward pass + adjoint kernel per direction)."""
    import torch  # noqa: F401  (device plumbing only)

    import sagittasbr_amd as S
    from sagittasbr_amd import workloads as W

    wl = W.c5()
    log(wl["desc"])
    ctx = S.Context(0)
    v, f = wl["meshes"][0]
    mid = ctx.mesh(v, f)
    th, ph = wl["directions"]
    jb = wl["job"]
    jobs = S.direction_jobs(mid, th, ph, jb["freq"], jb["grid_size"], jb["n_x"], jb["n_y"], jb["max_bounces"])
    plan = ctx.plan(jobs, S.SBR_MIXED, S.SBR_BATCHED)
    tuning = plan.autotune(sample_jobs=0, repeats=1)
    log(f"tuning {tuning}")
    ms, _ = timed(plan, max(args.steps, 1), args.warmup)
    z, sig, db = plan.results()
    plan.close()
    w = np.ones(len(th)) / len(th)
    ctx.adjoint(mid, jobs, w)  # warm-up (FP64 tree, device copies of the vertices)
    t0 = time.perf_counter()
    sg, dxyz, dsdf = ctx.adjoint(mid, jobs, w)
    adj_s = time.perf_counter() - t0
    ctx.close()
    log(f"forward {ms:.2f} ms, adjoint {adj_s:.3f} s")
    n = len(th)
    return {
        "metric": "angles/s (batched mixed-precision sweep, <=0.1 dB vs FP64), 1 GPU",
        "value": round(n / (ms / 1000.0), 2), "unit": "angles/s", "n_gpus": 1, "steps": max(args.steps, 1),
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic (procedural meshes)",
        "config": {"workload": wl["desc"], "method": "batched", "precision": "mixed", "tuning": tuning},
        "adjoint": {"seconds": round(adj_s, 4), "directions_per_s": round(n / adj_s, 2),
                    "gradient_values": int(dxyz.size), "nonzero_vertex_coordinates": int(np.count_nonzero(dxyz)),
                    "sigma_matches_forward": bool(np.allclose(sg, sig, rtol=1e-12)),
                    "dsigma_df_mean": float(np.mean(dsdf))},
    }


def other_workload(args):
    """c1 / c3 / c4: batched mixed sweep (autotuned, one persistent launch)
    against the plain per-angle FP32 method and the FP64 path, the last two
    on every stride-th angle (rates are per angle, accuracy per angle)."""
    import torch  # noqa: F401  (device plumbing only)

    import sagittasbr_amd as S
    from sagittasbr_amd import workloads as W

    t0 = time.perf_counter()
    wl = W.WORKLOADS[args.workload]()
    stride = args.stride or {"c1": 1, "c3": 31, "c4": 8}[args.workload]
    log(f"{wl['desc']}; generated in {time.perf_counter() - t0:.1f} s")
    ctx = S.Context(0)
    t0 = time.perf_counter()
    mids = [ctx.mesh(v, f) for v, f in wl["meshes"]]
    build_s = time.perf_counter() - t0
    log(f"{len(mids)} mesh(es) uploaded, FP32 BVH built in {build_s:.1f} s")

    def jobs(prec, step=1):
        out = []
        for mi, kw in wl["sweeps"]:
            js = S.sweep_jobs(S.SweepParams.make(**kw), mids[mi], prec)
            out.extend(js[k] for k in range(len(js)))
        out = out[::step]
        return (S.Job * len(out))(*out)

    jm = jobs(S.SBR_MIXED)
    n = len(jm)
    rays = sum(j.n_x * j.n_y for j in jm)
    plan = ctx.plan(jm, S.SBR_MIXED, S.SBR_BATCHED)
    log(f"{n} jobs, {rays:.3e} launched rays; autotuning")
    if args.tuning:
        tuning = json.loads(args.tuning)
        plan.tune(tuning["wg_per_cu"], tuning["lds_stack"], tuning["tile_rays"], tuning["use_graph"],
                  tuning.get("ray_pool", 0), tuning.get("traversal", 0))
    else:
        # the airframe's candidates differ by a few percent, and a short
        # sample run's tail favours small tiles that the full sweep does not
        # need: a larger sample keeps the choice close to the full sweep's
        tuning = plan.autotune(sample_jobs=min(n, 640 if args.workload == "c3" else 512), repeats=1)
    if args.workload == "c4":
        plan.tune(tuning["wg_per_cu"], tuning["lds_stack"], tuning["tile_rays"], 1, tuning["ray_pool"],
                  tuning["traversal"])
        tuning["use_graph"] = 1
    log(f"tuning {tuning}")
    steps = args.steps if args.steps else 1
    ms, kms = timed(plan, steps, args.warmup)
    log(f"batched mixed {ms:.1f} ms/step")
    z, sig, db = plan.results()
    own = plan.count()
    plan.close()

    j32 = jobs(S.SBR_FP32, stride)
    pa = ctx.plan(j32, S.SBR_FP32, S.SBR_PER_ANGLE)
    t_pa = pa.autotune(sample_jobs=min(len(j32), 64), repeats=1)
    pa_ms, _ = timed(pa, 1, 0)
    pa.close()
    per_angle_rate = len(j32) / (pa_ms / 1000.0)
    log(f"per-angle FP32 on {len(j32)} angles: {pa_ms:.1f} ms")

    j64 = jobs(S.SBR_FP64, stride)
    p64 = ctx.plan(j64, S.SBR_FP64, S.SBR_BATCHED)
    fp64_ms, _ = timed(p64, 1, 0)
    z64, s64, db64 = p64.results()
    cnt = p64.count()
    p64.close()
    log(f"FP64 on {len(j64)} angles: {fp64_ms:.1f} ms")
    derr = np.abs(db[::stride] - db64)
    strong = db64 > db64.max() - 30.0
    scale = n / len(j64)
    value = n / (ms / 1000.0)
    return {
        "metric": "angles/s (batched mixed-precision sweep, <=0.1 dB vs FP64), 1 GPU",
        "value": round(value, 2), "unit": "angles/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic (procedural meshes)",
        "config": {"workload": wl["desc"], "method": "batched", "precision": "mixed", "tuning": tuning,
                   "bvh_build_s": round(build_s, 2)},
        "ray_bounces_per_s_est": round(cnt["ray_bounces"] * scale / (ms / 1000.0), 1),
        "launched_rays": rays,
        "per_angle_fp32_angles_per_s": round(per_angle_rate, 2), "per_angle_fp32_tuning": t_pa,
        "per_angle_sample": f"every {stride}th angle ({len(j32)})",
        "batched_over_per_angle": round(value / per_angle_rate, 3),
        "fp64_angles_per_s": round(len(j64) / (fp64_ms / 1000.0), 2),
        "accuracy_vs_fp64_db": {"angles_compared": len(j64), "max": float(derr.max()),
                                "median": float(np.median(derr)),
                                "max_within_30dB_of_peak": float(derr[strong].max()),
                                "frac_angles_le_0.1dB": float(np.mean(derr <= 0.1))},
        "undecided_rays": own["retraced"],
        "full_fp64_queries": own["fp64_queries"],
        "counters_fp64_sample": cnt,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the sweep would shard by angle "
              "tiles across GPUs (not built: one GPU is measured)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.workload != "c2":
        if args.profile_run:
            raise SystemExit("--profile-run applies to the c2 workload")
        print(json.dumps(adjoint_workload(args) if args.workload == "c5" else other_workload(args)))
        return
    import torch  # device plumbing only: the sweep runs in libsbr.so

    import sagittasbr_amd as S
    from sagittasbr_amd import meshes as M

    prec = {"fp32": S.SBR_FP32, "fp64": S.SBR_FP64, "mixed": S.SBR_MIXED}[args.precision]
    verts, faces = M.trihedral_duct_scene()
    params = S.SweepParams.make(**WORKLOAD)
    ctx = S.Context(0)
    mesh = ctx.mesh(verts, faces)
    jobs = S.sweep_jobs(params, mesh, prec)
    plan = ctx.plan(jobs, prec, S.SBR_BATCHED)
    n_angles = params.n_angles
    if args.tuning:
        t = json.loads(args.tuning)
        plan.tune(t["wg_per_cu"], t["lds_stack"], t["tile_rays"], t["use_graph"], t.get("ray_pool", 0),
                  t.get("traversal", 0))
        tuning = t
    elif not args.no_autotune:
        tuning = plan.autotune(sample_jobs=0, repeats=1)
    else:
        tuning = {"wg_per_cu": 0, "lds_stack": 0, "tile_rays": 0, "use_graph": 0, "ray_pool": 0}

    log(f"tuning {tuning}; timing {args.steps} steps")
    ms_per_step, kernel_ms = timed(plan, args.steps, args.warmup)
    log(f"{ms_per_step:.3f} ms/step")
    z, sigma, dbsm = plan.results()
    if args.profile_run:
        print(json.dumps({"profile_run": True, "ms_per_step": ms_per_step, "kernel_ms": kernel_ms}))
        return
    own = plan.count()

    # FP64 batched plan: accuracy reference and the reference traversal's
    # node visits / triangle tests (bit-exact with the FP64 oracle), which
    # give the algorithmic bytes
    log("fp64 batched plan")
    jobs64 = S.sweep_jobs(params, mesh, S.SBR_FP64)
    p64 = ctx.plan(jobs64, S.SBR_FP64, S.SBR_BATCHED)
    t64 = p64.autotune(sample_jobs=0, repeats=1) if not args.no_autotune else None
    fp64_ms, _ = timed(p64, 2, 1)
    z64, s64, db64 = p64.results()
    cnt = p64.count()
    derr = np.abs(dbsm - db64)
    strong = db64 > db64.max() - 30.0
    alg_bytes = cnt["node_visits"] * NODE_BYTES + cnt["tri_tests"] * TRI_BYTES + OUT_BYTES * n_angles
    # the trace kernel is the only kernel of a run besides the output fills
    k_ms = float(np.mean(kernel_ms))
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9

    # FP32 batched and the plain per-angle FP32 method (one launch per angle,
    # its own autotuned launch shape), same GPU
    log("fp32 batched and per-angle plans")
    jobs32 = S.sweep_jobs(params, mesh, S.SBR_FP32)
    b32 = ctx.plan(jobs32, S.SBR_FP32, S.SBR_BATCHED)
    if not args.no_autotune:
        b32.autotune(sample_jobs=0, repeats=1)
    fp32_ms, _ = timed(b32, 2, 1)
    pa = ctx.plan(jobs32, S.SBR_FP32, S.SBR_PER_ANGLE)
    t_pa = pa.autotune(sample_jobs=512, repeats=1) if not args.no_autotune else None
    pa_ms, _ = timed(pa, 1, 1)
    per_angle_rate = n_angles / (pa_ms / 1000.0)

    value = n_angles / (ms_per_step / 1000.0)
    out = {
        "metric": "angles/s (batched mixed-precision sweep, <=0.1 dB vs FP64), 1 GPU",
        "value": round(value, 2),
        "unit": "angles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"mixed": "f32+f64", "fp32": "f32", "fp64": "f64"}[args.precision],
        "data": "synthetic (procedural trihedral + duct mesh)",
        "config": {"workload": "trihedral+duct multi-bounce sweep: 34,816 triangles, 10 GHz, 535x535 rays "
                               "(lambda/10 over 1.6 m), max 10 bounces, theta 0..90 x phi 0..90 (8,281 angles), "
                               "one persistent launch", "method": "batched", "precision": args.precision,
                   "tuning": tuning},
        "ray_bounces_per_s": round(cnt["ray_bounces"] / (ms_per_step / 1000.0), 1),
        "bvh_queries_per_s": round(cnt["queries"] / (ms_per_step / 1000.0), 1),
        "per_angle_fp32_angles_per_s": round(per_angle_rate, 2),
        "per_angle_fp32_tuning": t_pa,
        "batched_over_per_angle": round(value / per_angle_rate, 3),
        "batched_fp32_angles_per_s": round(n_angles / (fp32_ms / 1000.0), 2),
        "batched_fp64_angles_per_s": round(n_angles / (fp64_ms / 1000.0), 2),
        "accuracy_vs_fp64_db": {"max": float(derr.max()), "p99": float(np.percentile(derr, 99)),
                                "median": float(np.median(derr)),
                                "max_within_30dB_of_peak": float(derr[strong].max()),
                                "frac_angles_le_0.1dB": float(np.mean(derr <= 0.1))},
        "counters_fp64_traversal": cnt,
        "undecided_rays": own["retraced"],
        "full_fp64_queries": own["fp64_queries"],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": "trace_kernel", "kernel_ms": round(k_ms, 3), "algorithmic_bytes": int(alg_bytes)},
    }
    if not args.no_pmc:
        log("rocprofv3 counter passes")
        pm = pmc_measure(tuning)
        if "error" not in pm:
            out["roofline"]["traffic"] = int(pm["read_bytes"] + pm["write_bytes"])
            out["roofline"]["traffic_read"] = int(pm["read_bytes"])
            out["roofline"]["traffic_write"] = int(pm["write_bytes"])
            out["roofline"]["hbm_gbs"] = round((pm["read_bytes"] + pm["write_bytes"]) / (k_ms * 1e-3) / 1e9, 2)
            out["valu"] = {k: round(pm[k], 3) for k in ("valu_busy_pct", "valu_utilization_pct",
                                                       "mean_occupancy_waves_per_cu")}
            out["pmc_scale"] = {"fetch": round(pm["fetch_scale"], 4), "write": round(pm["write_scale"], 4)}
        else:
            out["pmc_error"] = pm["error"]
    if not args.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(verts, faces, WORKLOAD, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
