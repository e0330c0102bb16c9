"""Per-ray parity helpers: compare GPU per-ray records with the oracle and
classify the differences as edge/grazing ties.

A ray counts as a tie when its FP64 oracle outcome (hit ids per bounce) is
not stable under perturbations of the size of FP32 rounding: the origin is
moved by a few 1e-7 of the launch distance and the direction by a few 1e-7
rad. Such rays sit on a triangle edge, a silhouette or a grazing reflection,
where the reference's own FP32 and FP64 builds can disagree as well.
"""
from __future__ import annotations

import numpy as np

import oracle_lib as O


def ray_origin_dir(p: O.Params, angle: int, ray: int):
    fr, _ = O.frame(p, angle, precision=1)
    llc, U, V, d = fr
    i, j = ray % p.n_x, ray // p.n_x
    u = (i + 0.5) / p.n_x
    v = (j + 0.5) / p.n_y
    return llc + u * U + v * V, d


def outcome(ids_row, count):
    return tuple(int(x) for x in ids_row[:count])


def is_tie(scene64: O.Scene, p: O.Params, angle: int, ray: int, n_jitter=24, seed=0):
    """True when the FP64 outcome of the ray changes under FP32-sized
    perturbations of its origin and direction."""
    o, d = ray_origin_dir(p, angle, ray)
    scale = float(np.linalg.norm(o))
    rng = np.random.default_rng(seed + ray)
    base = scene64.trace_one(o, d, p.max_bounces)
    ref = outcome(base[1], base[0])
    for _ in range(n_jitter):
        amp = 10.0 ** rng.uniform(-8, -6)
        oj = o + rng.normal(size=3) * amp * scale
        dj = d + rng.normal(size=3) * amp
        dj /= np.linalg.norm(dj)
        n, ids, _ = scene64.trace_one(oj, dj, p.max_bounces)
        if outcome(ids, n) != ref:
            return True
    return False


def mismatches(a_ids, a_cnt, b_ids, b_cnt):
    """Indices of rays whose bounce count or hit ids differ."""
    diff = a_cnt != b_cnt
    diff |= np.any(a_ids != b_ids, axis=1)
    return np.nonzero(diff)[0]


def classify(scene64, p, angle, rays, gpu_ids, gpu_cnt, limit=400):
    """Returns (n_ties, n_real, real_rays) among the first ``limit`` rays."""
    ties, real = 0, []
    for r in rays[:limit]:
        if is_tie(scene64, p, angle, int(r)):
            ties += 1
        else:
            real.append(int(r))
    return ties, len(real), real
