"""Named strip-PET workloads: scanner geometry and seeded phantoms.

Geometry: strips at y = +-0.43 m, length 0.5 m, sigma_z = 0.01 m,
sigma_dl = 0.04 m.  Phantom regions are (x, y, a, b, angle_deg, intensity),
earlier regions on top (src/2d/geometry/phantom.h:193-203).
"""
import numpy as np

from .petstrip import Scanner, phantom_events

R, L, SIGMA_Z, SIGMA_DL = 0.43, 0.5, 0.01, 0.04

# three ellipses: cold circle, hot circle, elliptical background
PHANTOM_3 = [
    (-0.08, 0.03, 0.03, 0.03, 0.0, 0.0),
    (0.08, -0.02, 0.04, 0.04, 0.0, 4.0),
    (0.0, 0.0, 0.20, 0.15, 0.0, 1.0),
]


def shepp6(seed=7):
    """Six-ellipse Shepp-like phantom, intensities 0, 0.2, 1, 1.5, 2, 4,
    semi-axes 0.02-0.22 m, rotations drawn from `seed`."""
    rot = np.random.default_rng(seed).uniform(0.0, 180.0, 6)
    return [
        (0.05, 0.06, 0.025, 0.020, rot[0], 0.0),
        (-0.10, -0.04, 0.060, 0.040, rot[1], 4.0),
        (0.10, -0.06, 0.050, 0.030, rot[2], 2.0),
        (-0.06, 0.08, 0.070, 0.045, rot[3], 1.5),
        (0.00, 0.00, 0.150, 0.110, rot[4], 0.2),
        (0.00, 0.00, 0.220, 0.170, rot[5], 1.0),
    ]


def shepp10(seed=11):
    """Ten-ellipse phantom for the 0.5 mm workload."""
    rng = np.random.default_rng(seed)
    rot = rng.uniform(0.0, 180.0, 10)
    base = [
        (0.02, 0.07, 0.020, 0.020, 0.0), (-0.12, -0.05, 0.030, 0.020, 4.0),
        (0.12, 0.04, 0.025, 0.025, 3.0), (-0.05, -0.09, 0.040, 0.030, 0.5),
        (0.07, -0.08, 0.050, 0.030, 2.0), (-0.10, 0.07, 0.045, 0.035, 1.5),
        (0.15, -0.01, 0.030, 0.060, 2.5), (-0.01, 0.00, 0.090, 0.060, 0.2),
        (0.00, 0.00, 0.160, 0.120, 1.2), (0.00, 0.00, 0.220, 0.170, 1.0),
    ]
    return [(x, y, a, b, rot[i], c) for i, (x, y, a, b, c) in enumerate(base)]


WORKLOADS = {
    # name: (n_pixels, pixel_size, n_events, iterations, phantom, seed)
    "strip-160px-4mm-1e6ev": (160, 0.004, 1_000_000, 10, PHANTOM_3, 1),
    "strip-512px-1mm-2e7ev": (512, 0.001, 20_000_000, 10, shepp6(), 2),
    "strip-256px-2mm-5e6ev": (256, 0.002, 5_000_000, 1, PHANTOM_3, 4),
    "strip-1024px-0.5mm-1e8ev": (1024, 0.0005, 100_000_000, 5, shepp10(), 5),
}


def workload_scanner(name, sigma_z=SIGMA_Z, sigma_dl=SIGMA_DL):
    n_pix, size = WORKLOADS[name][:2]
    return Scanner(R, L, n_pix, n_pix, size, size, sigma_z, sigma_dl)


def expected_image(scanner, regions, n_events, sensitivity, sub=4):
    """The expected image of detected emissions of a phantom: per pixel the
    intensity of the top region (earlier regions on top, the generator's
    non-additive rule) averaged over sub x sub points of the pixel, times
    the detection probability (the scanner's sensitivity map at the pixel
    centres, Scanner::sensitivity, src/2d/strip/scanner.h:147-171, which is
    the fraction of emission angles seen by both strips), scaled to sum to
    n_events.  Rows run from the top (y = +H ph / 2), columns from the left
    (z = -W pw / 2), as the images of Reconstruction.  Regions are
    (x, y, a, b, angle_deg, intensity) ellipses, x along z."""
    W, H = scanner.n_z_pixels, scanner.n_y_pixels
    pw, ph = scanner.pixel_width, scanner.pixel_height
    off = (np.arange(sub) + 0.5) / sub
    z = (-W * pw / 2 + scanner.center_z) + (np.arange(W)[:, None] + off[None, :]).ravel() * pw
    y = (H * ph / 2 + scanner.center_y) - (np.arange(H)[:, None] + off[None, :]).ravel() * ph
    zz, yy = np.meshgrid(z, y)
    val = np.zeros(zz.shape)
    free = np.ones(zz.shape, bool)
    for x0, y0, a, b, ang, inten in regions:
        c, s_ = np.cos(np.radians(ang)), np.sin(np.radians(ang))
        rx, ry = zz - x0, yy - y0
        ex, ey = c * rx + s_ * ry, -s_ * rx + c * ry
        inside = (ex / a) ** 2 + (ey / b) ** 2 <= 1.0
        val[inside & free] = inten
        free &= ~inside
    img = val.reshape(H, sub, W, sub).mean(axis=(1, 3)) * np.asarray(sensitivity, np.float64)
    return (img * (n_events / img.sum())).astype(np.float32)


def workload_events(name, n=None, seed=None):
    n_pix, size, n_ev, _, phantom, s = WORKLOADS[name]
    sc = workload_scanner(name)
    return phantom_events(sc, phantom, n_ev if n is None else n,
                          seed=s if seed is None else seed)


# 3D hybrid workloads: one ring of detectors (hybrid_geometry.ring_geometry),
# a voxel grid centred on the axis, cylinder responses
# (hybrid_geometry.cylinder_responses).
# name: (n_detectors, ring radius, detector width, n_pixels, pixel size,
#        n_planes, n_events, sigma_z, sigma_dl, length, seed)
HYBRID_WORKLOADS = {
    "hybrid-ring192-160px-4mm-128planes-1e6ev":
        (192, 0.43, 0.007, 160, 0.004, 128, 1_000_000, SIGMA_Z, SIGMA_DL, L, 1),
}


def hybrid_workload(name, n=None):
    """(geometry, lors, zzd, sigma_z, sigma_dl, length) of a named hybrid
    workload; n overrides the event count (a prefix of the same stream)."""
    from . import hybrid_geometry as hg
    n_det, radius, width, n_pix, ps, n_planes, n_ev, sz, sdl, length, seed = \
        HYBRID_WORKLOADS[name]
    g = hg.ring_geometry(n_detectors=n_det, ring_radius=radius, detector_width=width,
                         n_pixels=n_pix, pixel_size=ps, n_planes=n_planes)
    lors, zzd = hg.cylinder_responses(g, n_ev if n is None else n, sigma_z=sz,
                                      sigma_dl=sdl, length=length, seed=seed)
    return g, lors, zzd, sz, sdl, length
