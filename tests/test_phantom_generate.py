"""The phantom command's generator on host threads (the device runs the same
slot functions; tests/test_gpu_phantom.py compares the two, and
tests/test_gpu_phantom_statistics.py runs these checks on device-generated
events).

Semantics pinned against the reference's rules (src/2d/geometry/phantom.h,
rectangle.h, ellipse.h, src/2d/strip/scanner.h:107-113,195-267,
src/2d/strip/phantom_cmd.cpp:283-340) by statistics with explicit
tolerances: region weights (intensity x area, intensity for points),
non-additive rejection against additive overlap, uniform points in rotated
ellipses and in rectangles, the detection rule, the noise widths, the
PixelGrid images (row 0 at the bottom) and the emission / detected modes.
Plain numpy restatements of each rule are the checks; a plausible slip (area
weights ignored, rejection skipped, a swapped noise width) fails one of
them."""
import math

import numpy as np
import pytest

R, L, SZ, SDL = 0.43, 0.5, 0.01, 0.04
# where the generator runs: -1 host threads (these CPU tests), or a device
# index (the same checks on the device's events)
DEVICE = -1


def generate(pkg, *args, **kwargs):
    kwargs.setdefault("device", DEVICE)
    return pkg.phantom_generate(*args, **kwargs)


def scanner(pkg):
    return pkg.Scanner(R, L, 64, 64, 0.004, 0.004, SZ, SDL)


def ell(x, y, a, b, ang, i):
    return (0, x, y, a, b, ang, i)


def rect(x, y, a, b, i):
    return (1, x, y, a, b, 0.0, i)


def point(x, y, i):
    return (2, x, y, 0.0, 0.0, 0.0, i)


def test_detected_mode_equals_event_generator(pkg):
    """Ellipses, detected mode: the events of petstrip_phantom_events."""
    sc = scanner(pkg)
    regs6 = [(-0.08, 0.03, 0.03, 0.03, 0.0, 0.0), (0.0, 0.0, 0.2, 0.15, 30.0, 1.0)]
    a = pkg.phantom_events(sc, regs6, 30000, seed=9)
    b = generate(pkg, sc, [(0,) + r for r in regs6], 30000, seed=9, detected=True)
    assert b["responses"].shape == (30000, 3)
    assert np.array_equal(a.view(np.uint32), b["responses"].view(np.uint32))
    assert b["n_emitted"] >= 30000


def test_region_weights_and_kinds(pkg):
    """Emissions per region follow intensity x area (ellipse pi a b,
    rectangle 4 a b) and intensity for points; every emission of a region
    lies inside it; rectangles and ellipses are filled uniformly."""
    sc = scanner(pkg)
    regs = [ell(-0.06, 0.02, 0.03, 0.015, 40.0, 2.0), rect(0.05, -0.03, 0.02, 0.01, 3.0),
            point(0.0, 0.08, 0.002)]
    w = np.array([2.0 * math.pi * 0.03 * 0.015, 3.0 * 4 * 0.02 * 0.01, 0.002])
    n = 400000
    res = generate(pkg, sc, regs, n, seed=3, full=True)
    ev = res["events"].astype(np.float64)
    # emission mode: the events returned are the detected ones; detection
    # depends on the point, so compare weights on the emitted image instead
    g = pkg.phantom_grid(256, 256, 0.001)
    img = generate(pkg, sc, regs, n, seed=3, grid=g, images=True)["emitted"]
    assert img.sum() == n  # every emission lies inside the grid
    yy, xx = np.mgrid[0:256, 0:256]
    cx, cy = -0.128 + (xx + 0.5) * 0.001, -0.128 + (yy + 0.5) * 0.001
    c, s = math.cos(math.radians(40)), math.sin(math.radians(40))
    u, v = c * (cx + 0.06) + s * (cy - 0.02), -s * (cx + 0.06) + c * (cy - 0.02)
    # pixels whose centre lies within half a pixel diagonal of a region
    in_e = (u / 0.0308) ** 2 + (v / 0.0158) ** 2 <= 1
    in_r = (np.abs(cx - 0.05) <= 0.0211) & (np.abs(cy + 0.03) <= 0.0111)
    pt = (np.abs(cx - 0.0) < 0.001) & (np.abs(cy - 0.08) < 0.001)
    counts = np.array([img[in_e].sum(), img[in_r].sum(), img[pt].sum()], np.float64)
    assert counts.sum() == n
    want = n * w / w.sum()
    assert np.all(np.abs(counts - want) < 5 * np.sqrt(want) + 1)
    # uniform fill: interior pixels of the rectangle have equal expectations
    inner = (np.abs(cx - 0.05) <= 0.018) & (np.abs(cy + 0.03) <= 0.008)
    vals = img[inner].astype(np.float64)
    assert abs(vals.std() / vals.mean() - 1 / math.sqrt(vals.mean())) < 0.1
    # detected events: their exact responses follow the emission points
    t = ev[:, 2] / ev[:, 3]
    ex = res["exact"].astype(np.float64)
    assert np.allclose(ex[:, 0], ev[:, 0] + (R - ev[:, 1]) * t, atol=2e-6)
    assert np.allclose(ex[:, 1], ev[:, 0] - (R + ev[:, 1]) * t, atol=2e-6)
    assert np.allclose(np.hypot(ev[:, 2], ev[:, 3]), 1, atol=1e-6)


def test_non_additive_rejection_and_additive_overlap(pkg):
    """A cold disc on top of a background: no emission inside the disc
    unless the phantom is additive."""
    sc = scanner(pkg)
    regs = [ell(0.0, 0.0, 0.03, 0.03, 0.0, 0.0), ell(0.0, 0.0, 0.1, 0.08, 0.0, 1.0)]
    g = pkg.phantom_grid(256, 256, 0.001)
    yy, xx = np.mgrid[0:256, 0:256]
    cx, cy = -0.128 + (xx + 0.5) * 0.001, -0.128 + (yy + 0.5) * 0.001
    disc = np.hypot(cx, cy) < 0.028
    n = 200000
    a = generate(pkg, sc, regs, n, seed=4, grid=g, images=True)["emitted"]
    b = generate(pkg, sc, regs, n, seed=4, grid=g, images=True, additive=True)["emitted"]
    assert a[disc].sum() == 0
    # additive: the background's share of the disc (area ratio)
    frac = b[disc].sum() / n
    want = (math.pi * 0.028 ** 2) / (math.pi * 0.1 * 0.08)
    assert abs(frac - want) < 0.05 * want


def test_detection_rule_and_noise(pkg):
    """Emission mode: detected iff both projected ends lie on the strips
    (acceptance against a numpy restatement from the emitted points); noise
    widths sigma_z, sigma_z, sigma_dl."""
    sc = scanner(pkg)
    regs = [ell(0.0, 0.0, 0.15, 0.1, 10.0, 1.0)]
    n = 300000
    res = generate(pkg, sc, regs, n, seed=5, full=True)
    assert res["n_emitted"] == n
    ev = res["events"].astype(np.float64)
    t = ev[:, 2] / ev[:, 3]
    zu, zd = ev[:, 0] + (R - ev[:, 1]) * t, ev[:, 0] - (R + ev[:, 1]) * t
    assert np.all((np.abs(zu) < L / 2 + 1e-6) & (np.abs(zd) < L / 2 + 1e-6))
    # acceptance: phi uniform in (-pi/2, pi/2), so the accepted fraction is
    # the mean over points of the accepted angle range / pi
    m = len(res["responses"])
    rng = np.random.default_rng(0)
    k = 200000
    rr, aa = np.sqrt(rng.random(k)), 2 * math.pi * rng.random(k)
    c, s = math.cos(math.radians(10)), math.sin(math.radians(10))
    ex_, ey_ = 0.15 * rr * np.cos(aa), 0.1 * rr * np.sin(aa)
    px, py = c * ex_ - s * ey_, s * ex_ + c * ey_
    phi = math.pi / 2 * (2 * rng.random(k) - 1)
    tt = np.tan(phi)
    ok = (np.abs(px + (R - py) * tt) < L / 2) & (np.abs(px - (R + py) * tt) < L / 2)
    p = ok.mean()
    assert abs(m / n - p) < 5 * math.sqrt(p * (1 - p) * (1 / n + 1 / k))
    d = res["responses"].astype(np.float64) - res["exact"].astype(np.float64)
    sd = d.std(axis=0)
    assert np.allclose(sd, [SZ, SZ, SDL], rtol=0.01)
    assert np.all(np.abs(d.mean(axis=0)) < 4 * np.array([SZ, SZ, SDL]) / math.sqrt(m))
    assert abs(np.corrcoef(d[:, 0], d[:, 1])[0, 1]) < 0.01


def test_images_and_tan_bins(pkg):
    """Detected image = histogram of the detected emission points; naive
    image = histogram of the centres from the responses with error
    (response.h:22-37); tan-bins map = the same centres by tan plane."""
    sc = scanner(pkg)
    regs = [ell(0.01, -0.02, 0.06, 0.04, 0.0, 1.0), point(-0.0513, 0.0517, 0.01)]
    g = pkg.phantom_grid(64, 48, 0.004, origin=(0.002, -0.004), tan_bins=8, max_tan=0.6)
    res = generate(pkg, sc, regs, 100000, seed=6, grid=g, images=True, full=True)
    ev = res["events"].astype(np.float64)

    def hist(x, y):
        col = np.floor((x - g.lower_left_x) / g.pixel_size)
        row = np.floor((y - g.lower_left_y) / g.pixel_size)
        ok = (col >= 0) & (col < 64) & (row >= 0) & (row < 48)
        h = np.zeros((48, 64), np.int64)
        np.add.at(h, (row[ok].astype(int), col[ok].astype(int)), 1)
        return h, ok, row, col

    # the events are stored in float: a few may sit across a pixel edge
    hd, _, _, _ = hist(ev[:, 0], ev[:, 1])
    assert np.abs(hd - res["detected"]).sum() <= 4
    assert hd.sum() > 1000
    r = res["responses"]
    t = (r[:, 0] - r[:, 1]) / np.float32(2 * R)
    y = np.float32(-0.5) * r[:, 2] / np.sqrt(np.float32(1) + t * t)
    z = np.float32(0.5) * (r[:, 0] + r[:, 1] + np.float32(2) * y * t)
    hn, ok, row, col = hist(z.astype(np.float64), y.astype(np.float64))
    # float rounding of the centre may move a handful across pixel edges
    assert np.abs(hn - res["naive"]).sum() <= 4
    plane = np.floor((t.astype(np.float64) + 0.6) / (1.2 / 8))
    okt = ok & (plane >= 0) & (plane < 8)
    tb = np.zeros((8, 48, 64), np.int64)
    np.add.at(tb, (plane[okt].astype(int), row[okt].astype(int), col[okt].astype(int)), 1)
    assert np.abs(tb - res["tan_bins_map"]).sum() <= 4
    assert res["emitted"].sum() == res["n_emitted"] - 0  # all inside the grid


def test_bad_arguments(pkg):
    sc = scanner(pkg)
    with pytest.raises(pkg.PetstripError):
        generate(pkg, sc, [(7, 0, 0, 0.1, 0.1, 0, 1)], 10)
    with pytest.raises(pkg.PetstripError):
        generate(pkg, sc, [rect(0, 0, -0.1, 0.1, 1)], 10)
    with pytest.raises(pkg.PetstripError):
        generate(pkg, sc, [point(0, 0, 0.0)], 10)


def test_expected_image_matches_generated_emissions(pkg):
    """workloads.expected_image (the phantom's expected image of detected
    emissions, which the calibration test holds fixed) against a histogram
    of the generator's own detected emission points on a 40 x 40 grid of
    16 mm: chi^2 per pixel about 1.  The same image with the rotations
    mirrored, or without the detection probability, is rejected."""
    import oracle_lib as ol
    regions = pkg.workloads.shepp6()
    sc = pkg.Scanner(0.43, 0.5, 40, 40, 0.016, 0.016, 0.01, 0.04)
    n = 400_000
    res = generate(pkg, sc, [(0,) + tuple(r) for r in regions], n, seed=17,
                   detected=True, full=True)
    pts = res["events"][:, :2].astype(np.float64)
    col = np.floor((pts[:, 0] + 0.32) / 0.016).astype(int)
    row = np.floor((0.32 - pts[:, 1]) / 0.016).astype(int)
    hist = np.zeros((40, 40))
    np.add.at(hist, (row, col), 1.0)
    sens = ol.sensitivity_f32(sc, 0)

    def chi2(img):
        m = img > 5
        return float(np.mean((hist[m] - img[m]) ** 2 / img[m]))
    good = chi2(pkg.expected_image(sc, regions, len(pts), sens, sub=8))
    mirrored = chi2(pkg.expected_image(sc, [r[:4] + (-r[4],) + r[5:] for r in regions],
                                       len(pts), sens, sub=8))
    flat = chi2(pkg.expected_image(sc, regions, len(pts), np.ones_like(sens), sub=8))
    print(f"chi2 per pixel: {good:.2f}, mirrored rotations {mirrored:.1f}, "
          f"no detection probability {flat:.1f}")
    assert good < 1.6
    assert mirrored > 5 * good and flat > 5 * good
