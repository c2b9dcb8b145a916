"""Writes the fixtures of the 3D hybrid command's --system paths: synthetic
system matrices and a scanner description file under
tests/golden/hybrid_system/, and tests/golden/hybrid_system.npz with what the
reference's own code (src/3d/hybrid/reconstruction.h, src/2d/barrel/
geometry_soa.h, sparse_matrix.h, scanner_builder.h), compiled from its
headers into oracle/_ref/libref_hybrid.so (ref_hybrid_run_system), makes of
them.

Run in a container that holds the reference tree:
    make -C oracle && python tests/golden/make_hybrid_system_golden.py

Inputs (synthetic, seeded): a ring of 24 detectors of radius 0.2 m around a
24 x 24 grid of 1 cm with 12 planes (hybrid_geometry.ring_geometry), 1500
responses of emissions within 1 cm of the central plane (sigma_z 6 mm, tan
theta at most 0.1), so
that no event reaches plane 0, where the reference's plane |iz - 6| lies
one plane past its weight and sensitivity arrays; six full-form matrices
(one per half plane) of hits falling off along each LOR with a few pixels
the geometry lacks, written with their LORs and pixels shuffled; a scanner
description file of the 24 detectors' corners.

Cases (arrays prefixed by the case name):
  am  the geometry with the six matrices as weight planes (the multi-plane
      geometry), the sensitivity volume calculated
  a1  the geometry with one matrix, sensitivity calculated
  b   the geometry of matrix 0 and the detector file, sensitivity calculated
  bi  the same with detectors 3 and 10 inactive, sensitivity one (with
      inactive detectors the reference's calculate_sensitivity reads
      pixel-info slots it never filled)
  bt  the geometry of a TOF matrix (magic TOFP, three TOF positions: each
      LOR holds its pixels up to three times) and the detector file
each with the weights (after normalisation), the sensitivity, the frame
events, the image after two iterations and, for b, bi and bt, the geometry
arrays.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "hybrid_system")
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle_lib as ol  # noqa: E402
from __graft_entry__ import load_package  # noqa: E402

SZ, SDL, LENGTH = 0.006, 0.05, 0.3
N_PLANES, N_MATRICES = 12, 6
N_ITER = 2


def inputs(pkg):
    hg = pkg.hybrid_geometry
    g = hg.ring_geometry(n_detectors=24, ring_radius=0.2, n_pixels=24, pixel_size=0.01,
                         n_planes=N_PLANES)
    lors, zzd = hg.cylinder_responses(g, 1500, sigma_z=SZ, sigma_dl=SDL, length=LENGTH,
                                      half_z=0.01, max_tan=0.1, seed=21)
    return g, lors, zzd


def main():
    if not ol.have_ref_hybrid():
        raise SystemExit("oracle/_ref/libref_hybrid.so missing: run make -C oracle")
    pkg = load_package()
    hg = pkg.hybrid_geometry
    g, lors, zzd = inputs(pkg)
    os.makedirs(OUT, exist_ok=True)
    mats = []
    for p in range(N_MATRICES):
        path = os.path.join(OUT, f"matrix_{p}.bin")
        hg.write_system_matrix(path, 24, 20000, 24,
                               hg.synthetic_system_matrix(g, 20000, seed=100 + p, plane=p))
        mats.append(path)
    tof = os.path.join(OUT, "matrix_tof.bin")
    hg.write_system_matrix(tof, 24, 20000, 24, hg.synthetic_tof_matrix(g, 20000, seed=200),
                           n_tof_positions=3)
    det = os.path.join(OUT, "detectors.txt")
    hg.write_detector_file(det, 24, 0.2)
    arrays = {"lors": lors, "zzd": zzd}
    cases = {
        "am": dict(mode=0, matrix_paths=mats, sens_mode=0),
        "a1": dict(mode=0, matrix_paths=mats[:1], sens_mode=0),
        "b": dict(mode=1, matrix_paths=mats[:1], detector_file=det, sens_mode=0,
                  pixel_size=0.01, n_planes=N_PLANES, z_left=-0.06),
        "bi": dict(mode=1, matrix_paths=mats[:1], detector_file=det, inactive=(3, 10),
                   sens_mode=1, pixel_size=0.01, n_planes=N_PLANES, z_left=-0.06),
        "bt": dict(mode=1, matrix_paths=[tof], detector_file=det, sens_mode=0,
                   pixel_size=0.01, n_planes=N_PLANES, z_left=-0.06),
    }
    for name, kw in cases.items():
        print(name, flush=True)
        r = ol.ref_hybrid_system(kw.pop("mode"), g, SZ, SDL, LENGTH, lors, zzd, N_ITER, **kw)
        ni = r["n_pixel_infos"]
        assert r["frames"][:, 8].min() >= 1, "an event reaches plane 0"
        arrays[f"{name}_weights"] = r["weights"]
        arrays[f"{name}_sensitivity"] = r["sensitivity"]
        arrays[f"{name}_frames"] = r["frames"]
        arrays[f"{name}_rho"] = r["rho"]
        arrays[f"{name}_n_pixel_infos"] = np.int64(ni)
        if name.startswith("b"):
            for k in ("segments", "begin", "end", "pixels", "positions"):
                arrays[f"{name}_{k}"] = r[k]
        print(f"{name}: {ni} pixel infos, {len(r['frames'])} frame events, "
              f"sum rho {r['rho'][-1].sum():.6g}")
    np.savez_compressed(os.path.join(HERE, "hybrid_system.npz"), **arrays)


if __name__ == "__main__":
    main()
