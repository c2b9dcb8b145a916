"""Writes the golden vectors under tests/golden/ from the reference's own
strip MLEM, compiled from its sources by oracle/Makefile into
oracle/_ref/libref_strip.so (a harness around
Strip::Reconstruction::operator(), src/2d/strip/reconstruction.h:108-154).

Run in a container that holds the reference tree:
    make -C oracle && python tests/golden/make_golden.py
Each fixture holds the input events and the reference's image after every
iteration (one OpenMP thread, so the accumulation order is the reference's
sequential event order), plus point values of the reference's kernel,
bounding box, response conversion, pixel mapping and sensitivity maps.
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
    # name: (R, L, n_y, n_z, pixel, sigma_z, sigma_dl, n_events, iterations, seed)
    "mlem_64px_8mm": (0.43, 0.5, 64, 64, 0.008, 0.01, 0.04, 8000, 3, 3),
    "mlem_41px_12mm": (0.43, 0.5, 41, 41, 0.0125, 0.01, 0.04, 4000, 2, 5),
    "mlem_48px_wide_sigma": (0.35, 0.4, 48, 48, 0.009, 0.015, 0.06, 4000, 2, 7),
}


def main():
    if not ol.have_ref():
        raise SystemExit("oracle/_ref/libref_strip.so missing: run make -C oracle "
                         "where the reference tree exists")
    pkg = load_package()
    for name, (R, L, ny, nz, ps, sz, sdl, n, it, seed) in CASES.items():
        sc = pkg.Scanner(R, L, ny, nz, ps, ps, sz, sdl)
        ev = pkg.phantom_events(sc, pkg.workloads.PHANTOM_3, n, seed=seed)
        rho = ol.ref_mlem(sc, ev, it, n_threads=1)
        sens0 = np.zeros(ny * nz, np.float32)
        sens1 = np.zeros(ny * nz, np.float32)
        r = ol.ref()
        r.ref_strip_sensitivity(R, L, ny, nz, ps, sz, sdl, ol.P(sens0))
        r.ref_strip_pixel_sensitivity(R, L, ny, nz, ps, sz, sdl, ol.P(sens1))
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"),
            scanner=np.array([R, L, ny, nz, ps, sz, sdl], np.float64),
            events=ev, rho=rho, sensitivity=sens0.reshape(ny, nz),
            pixel_sensitivity=sens1.reshape(ny, nz))
        print(name, ev.shape, rho.shape)

    # point values of the reference's float kernel, box, response and pixel map
    rng = np.random.default_rng(11)
    r = ol.ref()
    k_in = np.stack([rng.uniform(-0.3, 0.3, 64), rng.uniform(-0.7, 0.7, 64),
                     rng.uniform(-0.03, 0.03, 64), rng.uniform(-0.06, 0.06, 64)], 1)
    k_out = []
    for y, t, dz, dy in k_in:
        t32 = np.float32(t)
        sec = float(np.sqrt(np.float32(1) + t32 * t32, dtype=np.float32))
        k_out.append([sec, r.ref_strip_kernel_f32(0.01, 0.04, y, float(t32), sec, 0.43, dz, dy)])
    bb_t = rng.uniform(-0.8, 0.8, 64).astype(np.float32)
    bb_out = np.zeros((64, 6), np.float32)
    for i, t in enumerate(bb_t):
        r.ref_strip_ellipse_bb(0.01, 0.04, float(t), ol.P(bb_out[i]))
    tyz_in = np.stack([rng.uniform(-0.3, 0.3, 64), rng.uniform(-0.3, 0.3, 64),
                       rng.uniform(-0.5, 0.5, 64)], 1).astype(np.float32)
    tyz_out = np.zeros((64, 3), np.float32)
    for i in range(64):
        r.ref_strip_tan_y_z(0.43, ol.P(tyz_in[i]), ol.P(tyz_out[i]))
    pa_in = np.concatenate([rng.uniform(-0.3, 0.3, (60, 2)),
                            [[-0.256, 0.256], [0.256, -0.256], [-1e-7, 1e-7], [0.0, 0.0]]]
                           ).astype(np.float32)
    pa_out = np.zeros((64, 2), np.int32)
    for i, (z, y) in enumerate(pa_in):
        r.ref_strip_pixel_at(0.43, 0.5, 512, 512, 0.001, float(z), float(y), ol.P(pa_out[i]))
    np.savez_compressed(os.path.join(HERE, "points.npz"),
                        kernel_in=k_in, kernel_out=np.array(k_out), bb_tan=bb_t,
                        bb_out=bb_out, tyz_in=tyz_in, tyz_out=tyz_out,
                        pixel_at_in=pa_in, pixel_at_out=pa_out)
    print("points")


if __name__ == "__main__":
    main()
