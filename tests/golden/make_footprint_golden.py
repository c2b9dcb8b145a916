"""Writes tests/golden/footprints.npz: per-event footprints and pixel
centres taken from the reference's own code, compiled from its headers into
oracle/_ref/libref_strip.so (oracle/ref_driver.cpp).

Run in a container that holds the reference tree:
    make -C oracle && python tests/golden/make_footprint_golden.py

For the events of the three MLEM fixtures (tests/golden/mlem_*.npz):
  ell_<case>  the number of pixels Reconstruction::bb_update
              (src/2d/strip/reconstruction.h:220-298) writes for the event
              alone: the nonzero pixels of the reference's image after one
              iteration over that single event from rho = 1 (every pixel
              inside the 3-sigma ellipse gets K rho / denominator > 0)
  box_<case>  the clamped box bb_update scans, from the reference's
              Response::calculate_tan_y_z, GaussianKernel::ellipse_bb and
              Scanner::pixel_at (response.h:22-37, gaussian_kernel.h:
              108-128, scanner.h:142-144), with n_pixels_in_line and
              Pixel::clamp as written there (reconstruction.h:216-246,
              pixel.h:51-54)
  centre_<case> Scanner::pixel_center (scanner.h:136-140) in float, (z, y),
              for every pixel of the image
"""
import ctypes as C
import glob
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as ol  # noqa: E402

N_EVENTS = 600


def reference_box(r, R, L, ny, nz, ps, sz, sdl, ev):
    tyz = np.zeros(3, np.float32)
    r.ref_strip_tan_y_z(R, ev.ctypes.data, tyz.ctypes.data)
    t, y, z = tyz
    bb = np.zeros(6, np.float32)
    r.ref_strip_ellipse_bb(sz, sdl, float(t), bb.ctypes.data)
    px = np.zeros(2, np.int32)
    r.ref_strip_pixel_at(R, L, ny, nz, ps, z, y, px.ctypes.data)
    # n_pixels_in_line: int(length / pixel_size + 0.5) in float
    hw = int(np.float32(np.float32(bb[5]) / np.float32(ps)) + np.float32(0.5))
    hh = int(np.float32(np.float32(bb[4]) / np.float32(ps)) + np.float32(0.5))

    def s16(v):
        return int(np.int32(v).astype(np.int16))

    def clamp(v, hi):
        return min(hi, max(0, v))
    x0, x1 = clamp(s16(px[0] - hw), nz - 1), clamp(s16(px[0] + hw), nz - 1)
    y0, y1 = clamp(s16(px[1] - hh), ny - 1), clamp(s16(px[1] + hh), ny - 1)
    return max(0, x1 - x0) * max(0, y1 - y0) if (x1 > x0 and y1 > y0) else 0


def main():
    if not ol.have_ref():
        raise SystemExit("oracle/_ref/libref_strip.so missing")
    r = ol.ref()
    out = {}
    for path in sorted(glob.glob(os.path.join(HERE, "mlem_*.npz"))):
        f = np.load(path)
        name = os.path.basename(path)[len("mlem_"):-len(".npz")]
        R, L, ny, nz, ps, sz, sdl = f["scanner"]
        ny, nz = int(ny), int(nz)
        sc = ol.SimpleScanner(R, L, ny, nz, ps, ps, sz, sdl)
        evs = f["events"][:N_EVENTS]
        ell = np.zeros(len(evs), np.int32)
        box = np.zeros(len(evs), np.int32)
        for k, ev in enumerate(evs):
            e = np.ascontiguousarray(ev, np.float32)
            img = ol.ref_mlem(sc, e[None, :], 1)[0]
            ell[k] = int(np.count_nonzero(img))
            box[k] = reference_box(r, R, L, ny, nz, ps, sz, sdl, e)
        cen = np.zeros((ny, nz, 2), np.float32)
        for y in range(ny):
            for x in range(nz):
                o = np.zeros(2, np.float32)
                r.ref_strip_pixel_center(R, L, ny, nz, ps, x, y, o.ctypes.data)
                cen[y, x] = o
        out[f"ell_{name}"], out[f"box_{name}"], out[f"centre_{name}"] = ell, box, cen
        print(f"{name}: {len(evs)} events, mean box {box.mean():.1f}, mean ellipse "
              f"{ell.mean():.1f}")
    np.savez_compressed(os.path.join(HERE, "footprints.npz"), **out)


if __name__ == "__main__":
    main()
