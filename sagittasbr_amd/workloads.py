"""The benchmark configurations of the SBR ray path as job lists.

Each workload is a set of meshes plus sweeps over them, in the reference's
own terms (config.txt keys, include/app/config.hpp:7-27): a square launch
grid of n_x * n_y rays over grid_size, centred on the origin and facing the
radar, so grid_size is twice the body's bounding radius and the ray spacing
grid_size / n_x is the stated fraction of a wavelength.

  c1  PEC sphere, icosphere 5 (20,480 triangles), 3 GHz, lambda/10, 5 bounces,
      theta = 90, phi = 0..359 (360 angles)
  c2  trihedral (3 x 64 x 64 x 2) + open duct (5 x 32 x 32 x 2), 34,816
      triangles, 10 GHz, lambda/10, 10 bounces, theta, phi = 0..90 (8,281)
  c3  procedural airframe (~3.2M triangles), f = 1, 3, 10 GHz, lambda/8,
      8 bounces, theta = 60..120 step 2 x phi = 0..359 (11,160 per f)
  c4  64 seeded random bodies (5k-400k triangles each), own f ~ U[1, 10] GHz,
      lambda/10, 6 bounces, theta = 90, phi = 0..359 (23,040 angles)
  c5  SH-perturbed sphere (icosphere 7, 327,680 triangles), 3 GHz, lambda/10,
      4 bounces, 64 directions uniform on the sphere (the adjoint workload)
"""
from __future__ import annotations

import math

from . import meshes as M

C0 = 299792458.0


def _grid(radius, freq, frac):
    """grid_size covering a body of bounding radius `radius` and the ray count
    per side giving a spacing of at most lambda / frac."""
    g = 2.0 * radius * 1.01
    n = int(math.ceil(g / (C0 / freq / frac)))
    return g, n


def c1():
    v, f = M.icosphere(5)
    kw = dict(theta_start=90.0, theta_end=90.0, theta_samples=1, phi_start=0.0, phi_end=359.0, phi_samples=360,
              grid_size=2.5, n_x=250, n_y=250, max_bounces=5, freq=3e9)
    return dict(name="c1", meshes=[(v, f)], sweeps=[(0, kw)],
                desc="sphere r=1 m (icosphere 5, 20,480 triangles), 3 GHz, 250x250 rays over 2.5 m (lambda/10), "
                     "max 5 bounces, theta 90, phi 0..359 (360 angles)")


def c2():
    v, f = M.trihedral_duct_scene()
    kw = dict(theta_start=0.0, theta_end=90.0, theta_samples=91, phi_start=0.0, phi_end=90.0, phi_samples=91,
              grid_size=1.6, n_x=535, n_y=535, max_bounces=10, freq=10e9)
    return dict(name="c2", meshes=[(v, f)], sweeps=[(0, kw)],
                desc="trihedral+duct (34,816 triangles), 10 GHz, 535x535 rays over 1.6 m (lambda/10), max 10 "
                     "bounces, theta 0..90 x phi 0..90 (8,281 angles)")


def c3(freqs=(1e9, 3e9, 10e9)):
    v, f = M.airframe(1)
    r = M.bounding_radius(v)
    sweeps = []
    for fr in freqs:
        g, n = _grid(r, fr, 8.0)
        sweeps.append((0, dict(theta_start=60.0, theta_end=120.0, theta_samples=31, phi_start=0.0, phi_end=359.0,
                               phi_samples=360, grid_size=g, n_x=n, n_y=n, max_bounces=8, freq=fr)))
    return dict(name="c3", meshes=[(v, f)], sweeps=sweeps,
                desc=f"airframe ({len(f):,} triangles), f = {', '.join(f'{x / 1e9:g}' for x in freqs)} GHz, "
                     f"lambda/8 over {2.02 * r:.2f} m, max 8 bounces, theta 60..120 step 2 x phi 0..359 "
                     f"({11160 * len(freqs):,} angles)")


def c4(n_bodies=64):
    bodies = M.random_bodies(2, n_bodies)
    meshes, sweeps = [], []
    for k, b in enumerate(bodies):
        meshes.append((b["verts"], b["faces"]))
        g, n = _grid(M.bounding_radius(b["verts"]), b["freq"], 10.0)
        sweeps.append((k, dict(theta_start=90.0, theta_end=90.0, theta_samples=1, phi_start=0.0, phi_end=359.0,
                               phi_samples=360, grid_size=g, n_x=n, n_y=n, max_bounces=6, freq=b["freq"])))
    ntri = sum(len(m[1]) for m in meshes)
    return dict(name="c4", meshes=meshes, sweeps=sweeps,
                desc=f"{n_bodies} random bodies ({ntri:,} triangles in total, 5k-400k each), f ~ U[1,10] GHz, "
                     f"lambda/10, max 6 bounces, theta 90, phi 0..359 ({360 * n_bodies:,} angles)")


def c5_directions(n=64, seed=3):
    """n directions uniform on the sphere (theta, phi in degrees)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    u, w = rng.random(n), rng.random(n)
    return np.degrees(np.arccos(1 - 2 * u)), 360.0 * w


def c5():
    v, f, _ = M.sh_perturbed_sphere(7, seed=3)
    th, ph = c5_directions()
    return dict(name="c5", meshes=[(v, f)], directions=(th, ph),
                job=dict(freq=3e9, grid_size=2.5, n_x=250, n_y=250, max_bounces=4),
                desc=f"SH-perturbed sphere (icosphere 7, {len(f):,} triangles, 2 cm RMS), 3 GHz, 250x250 rays over "
                     "2.5 m (lambda/10), max 4 bounces, 64 directions uniform on the sphere")


WORKLOADS = {"c1": c1, "c2": c2, "c3": c3, "c4": c4, "c5": c5}
