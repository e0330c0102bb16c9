#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ (run in the build
container, where the reference tree and oracle/_ref exist).

* Data files the reference holds, copied verbatim as inputs/expected
  outputs: assets/test_models/box.obj and sphere.obj (meshes),
  assets/rcs_results.csv (sphere, 10 GHz, 10000^2 rays, 30 phi cuts) and
  assets/results_freq_scans.csv (sphere frequency scan, 2-column format).
* BVH vectors produced by the reference's own buildBvh
  (src/scene/bvhBuilder.cpp), compiled from the reference sources into
  oracle/_ref/ by oracle/Makefile, for a set of meshes, in FP32 and FP64.
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))
REF = "/root/reference"

import oracle_lib as O  # noqa: E402
from sagittasbr_amd import meshes as M  # noqa: E402


def main():
    for src, dst in [("assets/test_models/box.obj", "ref_box.obj"), ("assets/test_models/sphere.obj", "ref_sphere.obj"),
                     ("assets/rcs_results.csv", "ref_rcs_results_sphere_10GHz.csv"),
                     ("assets/results_freq_scans.csv", "ref_results_freq_scans.csv")]:
        shutil.copyfile(os.path.join(REF, src), os.path.join(HERE, dst))
    rng = np.random.default_rng(11)
    soup_v = rng.uniform(-1, 1, size=(3 * 300, 3)).astype(np.float32)
    soup_f = np.arange(3 * 300, dtype=np.uint32).reshape(-1, 3)
    cases = {
        "ref_box": M.read_obj(os.path.join(HERE, "ref_box.obj")),
        "ref_sphere": M.read_obj(os.path.join(HERE, "ref_sphere.obj")),
        "ico3": M.icosphere(3),
        "trihedral8": M.trihedral(0.5, 8),
        "random_soup": (soup_v, soup_f),
    }
    out = {}
    for name, (v, f) in cases.items():
        out[f"{name}_verts"] = v
        out[f"{name}_faces"] = f
        for prec in ("f32", "f64"):
            tris, bounds, links, root = O.ref_bvh(v, f, prec, parallel=False)
            out[f"{name}_{prec}_tris"] = tris
            out[f"{name}_{prec}_bounds"] = bounds
            out[f"{name}_{prec}_links"] = links
            out[f"{name}_{prec}_root"] = np.array([root])
    np.savez_compressed(os.path.join(HERE, "ref_bvh.npz"), **out)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
