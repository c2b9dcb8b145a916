"""Writes tests/golden/hybrid_*.npz from the reference's own 3D hybrid
reconstruction (PET3D::Hybrid::Reconstruction, src/3d/hybrid/
reconstruction.h), compiled from its sources by oracle/Makefile into
oracle/_ref/libref_hybrid.so.

Run in a container that holds the reference tree:
    make -C oracle && python tests/golden/make_hybrid_golden.py

Each fixture holds the geometry parameters (the geometry itself is rebuilt
by j-pet-mlem_amd/hybrid_geometry.ring_geometry, which is deterministic), the
responses, and the reference's normalised weights, sensitivity map, frame
events and images after every iteration (one thread: the reference's
sequential event order).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as ol  # noqa: E402
from __graft_entry__ import load_package  # noqa: E402

CASES = {
    # name: (ring: n_detectors, radius, width; grid: n_pixels, pixel size;
    #        sigma_z, sigma_dl, length, n responses, iterations, seed,
    #        sensitivity one)
    "hybrid_32px_48det": (48, 0.22, 0.007, 32, 0.0075, 0.01, 0.04, 0.3, 2000, 2, 5, False),
    "hybrid_24px_sens1": (32, 0.2, 0.009, 24, 0.01, 0.012, 0.05, 0.25, 1500, 2, 7, True),
}


def build(name):
    (n_det, radius, width, n_pix, ps, sz, sdl, length, n, n_iter, seed,
     sens_one) = CASES[name]
    hg = load_package().hybrid_geometry
    g = hg.ring_geometry(n_detectors=n_det, ring_radius=radius, detector_width=width,
                         n_pixels=n_pix, pixel_size=ps)
    lors, zzd = hg.cylinder_responses(g, n, sigma_z=sz, sigma_dl=sdl, length=length, seed=seed)
    return g, lors, zzd


def main():
    if not ol.have_ref_hybrid():
        sys.exit("oracle/_ref/libref_hybrid.so is missing: run make -C oracle with the "
                 "reference tree present")
    for name, c in CASES.items():
        g, lors, zzd = build(name)
        sz, sdl, length, n_iter, sens_one = c[5], c[6], c[7], c[9], c[11]
        r = ol.ref_hybrid(g, sz, sdl, length, lors, zzd, n_iter, sensitivity_one=sens_one)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), lors=lors, zzd=zzd,
                            weights=r["weights"], sensitivity=r["sensitivity"],
                            frames=r["frames"], rho=r["rho"])
        print(f"{name}: {g.n_pixel_infos} pixel infos, {len(r['frames'])} of {len(lors)} "
              f"events kept, rho sums {r['rho'].sum(axis=(1, 2, 3))}", flush=True)


if __name__ == "__main__":
    main()
