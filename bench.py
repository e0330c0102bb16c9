#!/usr/bin/env python3
"""Benchmark: the batched mixed-precision SBR sweep on one MI355X.

Default workload: the procedural airframe (3,244,032 triangles) swept at
f = 1, 3 and 10 GHz with lambda/8 ray spacing over its 22 m bounding square
(586^2 / 1,760^2 / 5,865^2 rays per angle), max 8 bounces, theta 60..120 in
2 degree steps x phi 0..359 (33,480 angles). The sweep's job list is split
into five interleaved slices (job j in slice j mod 5, so every slice holds
the same mix of frequencies and aspects); one step is one persistent launch
over one slice, and five consecutive steps sweep all 33,480 angles. Mesh,
BVHs and job lists are resident in HBM before the timed region; the launch
shape is autotuned before it on a strided sample of the jobs.

Prints one JSON line: angles/s of the whole sweep and ray-bounces/s, beside
(on slice 0, 6,696 angles) the plain per-angle FP32 method, batched FP32 and
FP64 on the same GPU, the dB error of the mixed and FP32 sweeps against FP64,
the roofline of the trace kernel (algorithmic VALU operations of the
reference traversal over the rays the launch traces, counted by an FP64
counting run with the timed run's culling, against the VALU peak; the same
over every ray of the culled rectangles beside it; HBM traffic, VALU and L2
activity from rocprofv3 counter passes of the same kernel in child
processes), and
the baseline leg: the CPU oracle on the host cores and the reference's own
HIP kernels (built from its sources) on a seeded sample of the same sweep.

--workload c1|c2|c4 runs the other sweep configurations the same way (one
slice: a step is the whole sweep); --workload c5 the adjoint workload.
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

METRIC = "angles/s & ray-bounces/s at ≤0.1 dB vs FP64, batched vs per-angle, 1 GPU"

# MI355X peaks. VALU: one full-rate operation per lane per clock on 256 CUs x
# 64 lanes at 2.4 GHz = 39.3 T lane-operations/s (the published FP64 vector
# peak, 78.6 TFLOP/s, counts an FMA as 2 flops; the FP32 vector peak,
# 157.3 TFLOP/s, adds packed FP32 issue, which no HIP code here uses).
VALU_PEAK_TOPS = 78.6 / 2.0
VALU_ISSUE_PEAK_G = 256 * 2.4  # wave64 full-rate VALU instructions per ns (4 SIMDs x 1 per 4 clocks per CU)
HBM_PEAK_GBS = 8000.0
# Algorithmic VALU operations of the reference traversal (rtKernels.cu) per
# unit of work, fused multiply-adds counted once, a division as one
# reciprocal: slab test of one box (rtKernels.cu:46-61) with the reciprocal
# direction formed once per query: 6 FMA, 6 min/max for the plane pairs, 4
# to combine, 2 clamps, 1 compare = 19; Moller-Trumbore (rtKernels.cu:10-44)
# with the edges formed per test: 45; the bounce update per query (normal,
# reflection, offset, path length; rtKernels.cu:131-152): 30.
OPS_BOX = 19
OPS_TRI = 45
OPS_QUERY = 30
COPY_BYTES = 1 << 30  # calibration copy: 1 GiB read + 1 GiB written per launch
DOMINANT = "trace_kernel"
SLICES = {"c3": 5}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c3", "c4", "c5"],
                    help="c3 (default): the airframe sweep; c1, c2, c4 the other sweeps, c5 the adjoint")
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the baseline leg (CPU and reference GPU)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU time of the baseline sample")
    ap.add_argument("--no-pmc", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 / FP32 / per-angle legs")
    ap.add_argument("--tuning", default="", help="JSON tuning to apply instead of autotuning")
    ap.add_argument("--profile-run", action="store_true",
                    help="only slice 0's batched sweep, --steps times (for rocprofv3 under a child process)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="with --profile-run: first three 1 GiB device copies (FETCH/WRITE_SIZE calibration)")
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
    del a, b


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def baseline_leg(wl, target_s, do_ref):
    """Baselines timed beside the measurement, never instead of it: the CPU
    oracle (a C++/OpenMP restatement of the reference path; the reference has
    no CPU ray path) on the host cores, and the reference's own HIP kernels
    (rtKernels.cu / poKernels.cu built from its sources, launched as runSweep
    launches them) on this GPU, each on a seeded sample with the same number
    of angles from every sweep of the workload (so the sample's rate
    estimates the whole job list's)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    rng = np.random.default_rng(7)
    sweeps = wl["sweeps"]
    orders = [rng.permutation(kw["theta_samples"] * kw["phi_samples"]).astype(np.int32) for _, kw in sweeps]
    out = {}
    # CPU: one angle per sweep to size the sample, then k per sweep
    scenes = {}
    t_one = 0.0
    for (mi, kw), order in zip(sweeps, orders):
        if mi not in scenes:
            v, f = wl["meshes"][mi]
            scenes[mi] = O.Scene(v, f, precision=0)
        t0 = time.perf_counter()
        scenes[mi].sweep(O.Params(**kw), order[:1])
        t_one += time.perf_counter() - t0
    k = int(max(1, min(min(len(o) for o in orders) - 1, target_s / max(t_one, 1e-6))))
    t0 = time.perf_counter()
    for (mi, kw), order in zip(sweeps, orders):
        scenes[mi].sweep(O.Params(**kw), order[1:1 + k])
    dt = time.perf_counter() - t0
    n = k * len(sweeps)
    out["cpu_baseline"] = {
        "value": round(n / dt, 4), "unit": "angles/s", "cores": int(O.lib().orc_num_threads()), "kind": "port",
        "cpu": cpu_model(),
        "sample": f"{k} seeded random angles of each of the {len(sweeps)} sweeps ({n} of "
                  f"{sum(len(o) for o in orders)}), FP32 oracle (reference arithmetic, same BVH), {dt:.1f} s"}
    del scenes
    if not do_ref:
        return out
    try:
        # reference kernels: eight seeded angles per sweep after one warm-up angle
        rt = po = 0.0
        nang = 0
        refs = {}
        for (mi, kw), order in zip(sweeps, orders):
            if mi not in refs:
                v, f = wl["meshes"][mi]
                refs[mi] = O.RefGpu(v, f, "f32")
            op = O.Params(**kw)
            refs[mi].sweep(op, order[-1:])
            _, (a, b) = refs[mi].sweep(op, order[-9:-1])
            rt += a
            po += b
            nang += 8
        for r in refs.values():
            r.close()
        out["ref_gpu_baseline"] = {
            "value": round(nang / ((rt + po) / 1000.0), 3), "unit": "angles/s",
            "ms_per_angle": round((rt + po) / nang, 3), "rt_kernel_ms": round(rt, 3), "po_kernel_ms": round(po, 3),
            "kind": "reference FP32 HIP build (launchRaysMultiBounce + integratePoMultiBounce, kernel time from "
                    "hipEvents as sweep.cu:562 reports it)",
            "sample": f"8 seeded random angles of each of the {len(sweeps)} sweeps"}
    except Exception as e:  # a baseline is reported beside the measurement, never instead of it
        out["ref_gpu_baseline"] = {"error": str(e)[:300]}
    return out


def _pmc_pass(counters, extra_args, tmp, name):
    """One rocprofv3 counter pass over a child bench process; returns
    {(kernel, counter): {dispatch: value}} summed over counter instances."""
    out = os.path.join(tmp, name)
    cmd = ["rocprofv3", "--pmc", *counters, "--output-format", "csv", "-d", out, "-o", name, "--",
           sys.executable, os.path.join(ROOT, "bench.py"), *extra_args]
    env = dict(os.environ, TMPDIR="/tmp")
    r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=240)
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


def _per_dispatch(res, kernel_sub, counter, top=0):
    """Mean over the kernel's dispatches (top > 0: over its `top` largest
    values -- the calibration copies, not the small host-to-device upload
    copies that run under the same kernel name)."""
    vals = [v for (k, c), d in res.items() if kernel_sub in k and c == counter for v in d.values()]
    if not vals:
        raise RuntimeError(f"no {counter} rows for {kernel_sub}")
    if top:
        vals = sorted(vals)[-top:]
    return float(np.mean(vals))


def pmc_measure(workload, tuning):
    """Counters of the trace kernel over slice 0, two passes (each a child
    process that first runs three 1 GiB device copies for calibration):
    FETCH_SIZE + VALU activity and instruction count; WRITE_SIZE + L2 hits
    and misses + occupancy. FETCH/WRITE_SIZE are scaled by the bytes a
    device copy of known size moves over what the same counter reads for it."""
    if shutil.which("rocprofv3") is None:
        return {"error": "rocprofv3 not found"}
    tmp = tempfile.mkdtemp(prefix="sbr_pmc_", dir="/tmp")
    prof = ["--profile-run", "--copy-calibration", "--workload", workload, "--steps", "2", "--tuning",
            json.dumps(tuning)]
    try:
        a = _pmc_pass(["FETCH_SIZE", "VALUBusy", "VALUUtilization", "SQ_INSTS_VALU"], prof, tmp, "passa")
        b = _pmc_pass(["WRITE_SIZE", "TCC_HIT_sum", "TCC_MISS_sum", "MeanOccupancyPerCU"], prof, tmp, "passb")
        f_b = _per_dispatch(a, DOMINANT, "FETCH_SIZE")
        w_b = _per_dispatch(b, DOMINANT, "WRITE_SIZE")
        f_c = _per_dispatch(a, "copyBuffer", "FETCH_SIZE", top=3)
        w_c = _per_dispatch(b, "copyBuffer", "WRITE_SIZE", top=3)
        # counters are in KiB per dispatch; scale each by copy bytes / copy reading
        fs = COPY_BYTES / (f_c * 1024.0)
        ws = COPY_BYTES / (w_c * 1024.0)
        return {"read_bytes": f_b * 1024.0 * fs, "write_bytes": w_b * 1024.0 * ws, "fetch_scale": fs,
                "write_scale": ws, "fetch_kib_raw": f_b, "write_kib_raw": w_b, "copy_fetch_kib_raw": f_c,
                "copy_write_kib_raw": w_c,
                "valu_busy_pct": _per_dispatch(a, DOMINANT, "VALUBusy"),
                "valu_utilization_pct": _per_dispatch(a, DOMINANT, "VALUUtilization"),
                "sq_insts_valu": _per_dispatch(a, DOMINANT, "SQ_INSTS_VALU"),
                "tcc_hit": _per_dispatch(b, DOMINANT, "TCC_HIT_sum"),
                "tcc_miss": _per_dispatch(b, DOMINANT, "TCC_MISS_sum"),
                "mean_occupancy_waves_per_cu": _per_dispatch(b, DOMINANT, "MeanOccupancyPerCU")}
    except Exception as e:  # counters are reported beside the measurement, never instead of it
        return {"error": str(e)[:300]}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def apply_tuning(plan, t):
    plan.tune(t["wg_per_cu"], t["lds_stack"], t["tile_rays"], t["use_graph"], t.get("ray_pool", 0),
              t.get("traversal", 0), t.get("top_levels", 0), t.get("node_width", 0), t.get("cull", 0))


def build_workload(S, wl, ctx):
    t0 = time.perf_counter()
    mids = [ctx.mesh(v, f) for v, f in wl["meshes"]]
    # the mixed mode also reads each mesh's FP64 tree: build it before timing
    for m in mids:
        ctx.mesh_info(m, S.SBR_FP64)
    return mids, time.perf_counter() - t0


def all_jobs(S, wl, mids, prec):
    out = []
    for mi, kw in wl["sweeps"]:
        js = S.sweep_jobs(S.SweepParams.make(**kw), mids[mi], prec)
        out.extend(js[k] for k in range(len(js)))
    return out


def job_array(S, jobs):
    return (S.Job * len(jobs))(*jobs)


def profile_run(args):
    """Child of the counter passes: slice 0 of the workload's batched mixed
    sweep with the given tuning, --steps launches."""
    import torch  # noqa: F401  (device plumbing only)

    import sagittasbr_amd as S
    from sagittasbr_amd import workloads as W

    if args.copy_calibration:
        copy_calibration()
    wl = W.WORKLOADS[args.workload]()
    ctx = S.Context(0)
    mids, _ = build_workload(S, wl, ctx)
    ns = SLICES.get(args.workload, 1)
    jm = all_jobs(S, wl, mids, S.SBR_MIXED)
    plan = ctx.plan(job_array(S, jm[0::ns]), S.SBR_MIXED, S.SBR_BATCHED)
    apply_tuning(plan, json.loads(args.tuning))
    ms = []
    for _ in range(max(args.steps, 1)):
        plan.run()
        plan.wait()
        ms.append(plan.last_ms())
    plan.close()
    ctx.close()
    print(json.dumps({"profile_run": True, "workload": args.workload, "slice_jobs": len(jm[0::ns]),
                      "kernel_ms": ms}))


def timed(plans, steps, warmup):
    """warmup untimed steps, then exactly `steps` timed ones; step s runs
    plans[s % len(plans)]. Returns wall ms per step and each step's
    trace-kernel milliseconds (hipEvents on the plan's stream). Every step
    ends with plan.wait(), a synchronisation of libsbr's stream, which is
    what brackets the timed region (torch's wheel carries a HIP runtime of
    its own; its synchronize is kept only as the harness's convention)."""
    import torch

    n = len(plans)
    for s in range(warmup):
        plans[s % n].run()
        plans[s % n].wait()
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for s in range(steps):
        p = plans[s % n]
        p.run()
        p.wait()
        kernel_ms.append(p.last_ms())
    torch.cuda.synchronize()
    return 1000.0 * (time.perf_counter() - t0) / steps, kernel_ms


def db_err(db, ref):
    d = np.abs(db - ref)
    strong = ref > ref.max() - 30.0
    return {"angles_compared": int(len(d)), "max": float(d.max()), "p99": float(np.percentile(d, 99)),
            "median": float(np.median(d)), "max_within_30dB_of_peak": float(d[strong].max()),
            "frac_angles_le_0.1dB": float(np.mean(d <= 0.1))}


def sweep_workload(args):
    """Batched mixed sweep (autotuned, one persistent launch per step over
    one slice of the job list), then the comparison legs on slice 0, the
    roofline, the counter passes and the baseline leg."""
    import torch  # noqa: F401  (device plumbing only)

    import sagittasbr_amd as S
    from sagittasbr_amd import workloads as W

    t0 = time.perf_counter()
    wl = W.WORKLOADS[args.workload]()
    log(f"{wl['desc']}; generated in {time.perf_counter() - t0:.1f} s")
    ctx = S.Context(0)
    mids, build_s = build_workload(S, wl, ctx)
    log(f"{len(mids)} mesh(es): FP32, four-wide and FP64 BVHs built and uploaded in {build_s:.1f} s")
    ns = SLICES.get(args.workload, 1)
    jm = all_jobs(S, wl, mids, S.SBR_MIXED)
    n = len(jm)
    slices = [job_array(S, jm[s::ns]) for s in range(ns)]
    plans = [ctx.plan(sl, S.SBR_MIXED, S.SBR_BATCHED) for sl in slices]
    rays = sum(j.n_x * j.n_y for j in jm)
    log(f"{n} jobs in {ns} slice(s), {rays:.3e} launched rays; autotuning")
    if args.tuning:
        tuning = json.loads(args.tuning)
    else:
        # a sample of slice 0; the airframe's candidates differ by a few
        # percent, and a short sample run's tail favours small tiles the
        # full sweep does not need: the sample is kept large
        # (the other sweeps are short enough to tune on all of their jobs:
        # a small sample's tail favours small tiles there too)
        sample = 640 if args.workload == "c3" else 0
        tuning = plans[0].autotune(sample_jobs=sample if sample < len(slices[0]) else 0, repeats=1)
    if args.workload == "c4":
        tuning["use_graph"] = 1
    for p in plans:
        apply_tuning(p, tuning)
    log(f"tuning {tuning}; timing {args.steps} steps after {args.warmup} warm-up steps")
    ms, kms = timed(plans, args.steps, args.warmup)
    log(f"batched mixed {ms:.2f} ms/step")
    sig = np.zeros(n)
    db = np.zeros(n)
    for s, p in enumerate(plans):
        _, sg, d = p.results()
        sig[s::ns] = sg
        db[s::ns] = d
    # trace-kernel time of slice 0's timed launches (the launch the
    # algorithmic counts and the counter passes cover)
    k0 = [m for s, m in enumerate(kms) if s % ns == 0]
    k0_ms = float(np.mean(k0)) if k0 else float(np.mean(kms))
    # angles swept in the timed steps over their wall time
    value = sum(len(slices[s % ns]) for s in range(args.steps)) / (ms * args.steps / 1000.0)
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "angles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic (procedural meshes)",
        "config": {"workload": wl["desc"], "method": "batched", "precision": "mixed", "slices": ns,
                   "step": (f"one persistent launch over one of {ns} interleaved slices of the {n} jobs "
                            f"(job j in slice j mod {ns}); {ns} steps sweep every angle once") if ns > 1
                   else f"one persistent launch over all {n} jobs",
                   "tuning": tuning, "bvh_build_s": round(build_s, 2)},
        "sweep_s": round(ms * ns / 1000.0, 3),
        "launched_rays": int(rays),
    }
    for p in plans[1:]:
        p.close()
    if not args.no_compare:
        n0 = len(slices[0])

        def same_jobs(prec):
            # slice 0's jobs with the angles of that precision's sweep_jobs
            # (FP32 sweeps form theta/phi in float, as the reference build does)
            jp = all_jobs(S, wl, mids, prec)
            return job_array(S, jp[0::ns])

        log(f"FP64 batched on slice 0 ({n0} jobs)")
        p64 = ctx.plan(same_jobs(S.SBR_FP64), S.SBR_FP64, S.SBR_BATCHED)
        apply_tuning(p64, dict(tuning, use_graph=0))
        fp64_ms, _ = timed([p64], 1, 0)
        _, _, db64 = p64.results()
        log("FP64 counting runs (reference traversal counters: all rays, then the rays a timed run traces)")
        cnt = p64.count()
        cnt_tr = p64.count(traced=True)
        p64.close()
        log("mixed counting run")
        own = plans[0].count()
        log("FP32 batched on slice 0")
        j32 = same_jobs(S.SBR_FP32)
        b32 = ctx.plan(j32, S.SBR_FP32, S.SBR_BATCHED)
        t32 = b32.autotune(sample_jobs=min(n0, 128), repeats=1)
        fp32_ms, _ = timed([b32], 1, 1)
        _, _, db32 = b32.results()
        b32.close()
        log("per-angle FP32 on slice 0")
        pa = ctx.plan(j32, S.SBR_FP32, S.SBR_PER_ANGLE)
        t_pa = pa.autotune(sample_jobs=min(n0, 64), repeats=1)
        pa_ms, _ = timed([pa], 1, 0)
        pa.close()
        per_angle_rate = n0 / (pa_ms / 1000.0)
        # the work of the rays the timed launch traces (the coverage masks
        # skip patches none of whose rays can reach the mesh)
        alg_ops = OPS_BOX * cnt_tr["node_visits"] + OPS_TRI * cnt_tr["tri_tests"] + OPS_QUERY * cnt_tr["queries"]
        alg_all = OPS_BOX * cnt["node_visits"] + OPS_TRI * cnt["tri_tests"] + OPS_QUERY * cnt["queries"]
        achieved = alg_ops / (k0_ms * 1e-3) / 1e12
        out.update({
            "ray_bounces_per_s": round(cnt["ray_bounces"] / (k0_ms * 1e-3), 1),
            "bvh_queries_per_s": round(cnt_tr["queries"] / (k0_ms * 1e-3), 1),
            "per_angle_fp32_angles_per_s": round(per_angle_rate, 2), "per_angle_fp32_tuning": t_pa,
            "batched_over_per_angle": round(value / per_angle_rate, 3),
            "batched_fp32_angles_per_s": round(n0 / (fp32_ms / 1000.0), 2), "batched_fp32_tuning": t32,
            "batched_fp64_angles_per_s": round(n0 / (fp64_ms / 1000.0), 2),
            "compare_set": f"slice 0: {n0} of {n} angles, every {ns}th job" if ns > 1 else f"all {n} angles",
            "accuracy_vs_fp64_db": db_err(db[0::ns], db64),
            "fp32_accuracy_vs_fp64_db": db_err(db32, db64),
            "counters_fp64_traversal_slice0": cnt,
            "counters_fp64_traversal_traced_slice0": cnt_tr,
            "undecided_rays_slice0": own["retraced"], "full_fp64_queries_slice0": own["fp64_queries"],
            "roofline": {"bound": "valu", "achieved": round(achieved, 4), "peak": VALU_PEAK_TOPS,
                         "unit": "Tops/s", "frac": round(achieved / VALU_PEAK_TOPS, 5), "traffic": None,
                         "kernel": "trace_kernel (mixed, persistent)", "kernel_ms": round(k0_ms, 3),
                         "algorithmic_ops": int(alg_ops),
                         "ops_model": f"{OPS_BOX} per box test x node visits + {OPS_TRI} per triangle test + "
                                      f"{OPS_QUERY} per query, reference traversal counts (FP64 counting run "
                                      f"over the rays a timed run of the same launch's jobs traces)",
                         "algorithmic_ops_all_rays": int(alg_all),
                         "frac_all_rays": round(alg_all / (k0_ms * 1e-3) / 1e12 / VALU_PEAK_TOPS, 5)},
        })
    plans[0].close()
    if not args.no_pmc and "roofline" in out:
        log("rocprofv3 counter passes")
        pm = pmc_measure(args.workload, tuning)
        if "error" not in pm:
            r = out["roofline"]
            r["traffic"] = int(pm["read_bytes"] + pm["write_bytes"])
            r["traffic_read"] = int(pm["read_bytes"])
            r["traffic_write"] = int(pm["write_bytes"])
            r["hbm_gbs"] = round((pm["read_bytes"] + pm["write_bytes"]) / (k0_ms * 1e-3) / 1e9, 2)
            r["hbm_frac"] = round(r["hbm_gbs"] / HBM_PEAK_GBS, 6)
            issue = pm["sq_insts_valu"] / (k0_ms * 1e-3) / 1e9
            out["valu"] = {"valu_busy_pct": round(pm["valu_busy_pct"], 3),
                           "valu_utilization_pct": round(pm["valu_utilization_pct"], 3),
                           "sq_insts_valu": int(pm["sq_insts_valu"]),
                           "issue_ginst_per_s": round(issue, 2),
                           "issue_frac": round(issue / VALU_ISSUE_PEAK_G, 4),
                           "executed_lane_ops_per_alg_op": round(
                               pm["sq_insts_valu"] * 64 * pm["valu_utilization_pct"] / 100.0 / out["roofline"][
                                   "algorithmic_ops"], 3),
                           "mean_occupancy_waves_per_cu": round(pm["mean_occupancy_waves_per_cu"], 3)}
            hm = pm["tcc_hit"] + pm["tcc_miss"]
            out["l2"] = {"requests": int(hm), "hit_pct": round(100.0 * pm["tcc_hit"] / max(hm, 1), 3),
                         "requests_per_s": round(hm / (k0_ms * 1e-3), 1)}
            out["pmc_scale"] = {"fetch": round(pm["fetch_scale"], 4), "write": round(pm["write_scale"], 4)}
            out["pmc_raw"] = {k: pm[k] for k in ("fetch_kib_raw", "write_kib_raw", "copy_fetch_kib_raw",
                                                 "copy_write_kib_raw", "sq_insts_valu", "valu_busy_pct",
                                                 "valu_utilization_pct", "tcc_hit", "tcc_miss",
                                                 "mean_occupancy_waves_per_cu")}
        else:
            out["pmc_error"] = pm["error"]
    ctx.close()
    if not args.no_cpu_baseline:
        log("baseline leg: CPU oracle and the reference's HIP kernels")
        out.update(baseline_leg(wl, args.cpu_seconds, do_ref=True))
    return out


def adjoint_workload(args):
    """c5: batched mixed forward of the 64 directions, then the frozen-path
    adjoint (recording forward pass + adjoint kernel)."""
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
    ms, _ = timed([plan], max(args.steps, 1), args.warmup)
    z, sig, db = plan.results()
    plan.close()
    w = np.ones(len(th)) / len(th)
    ctx.adjoint(mid, jobs, w)  # warm-up (device copies of the vertices)
    t0 = time.perf_counter()
    sg, dxyz, dsdf = ctx.adjoint(mid, jobs, w)
    adj_s = time.perf_counter() - t0
    ctx.close()
    log(f"forward {ms:.2f} ms, adjoint {adj_s:.3f} s")
    n = len(th)
    return {
        "metric": METRIC, "value": round(n / (ms / 1000.0), 2), "unit": "angles/s", "n_gpus": 1,
        "steps": max(args.steps, 1), "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic (procedural meshes)",
        "config": {"workload": wl["desc"], "method": "batched", "precision": "mixed", "tuning": tuning},
        "adjoint": {"seconds": round(adj_s, 4), "directions_per_s": round(n / adj_s, 2),
                    "gradient_values": int(dxyz.size), "nonzero_vertex_coordinates": int(np.count_nonzero(dxyz)),
                    "sigma_matches_forward": bool(np.allclose(sg, sig, rtol=1e-12)),
                    "dsigma_df_mean": float(np.mean(dsdf))},
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the sweep would shard by angle "
              "tiles across GPUs (not built: one GPU is measured)", file=sys.stderr)
        sys.exit(2)
    if args.profile_run:
        if not args.tuning:
            raise SystemExit("--profile-run needs --tuning")
        profile_run(args)
        return
    print(json.dumps(adjoint_workload(args) if args.workload == "c5" else sweep_workload(args)))


if __name__ == "__main__":
    main()
