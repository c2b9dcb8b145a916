"""Pins the CPU oracle (CPU, no GPU).

1. The reference's own unit-test goldens:
   gaussian_kernel_test.cpp:15-51, kernel_test.cpp:31-49 and 81-113,
   response_test.cpp:11-55, scanner_test.cpp:15-66 (src/2d/strip/).
2. Golden vectors written by tests/golden/make_golden.py from the reference
   compiled from its own sources: whole MLEM iterations must match bit for
   bit with one thread, point values of the kernel, box, response conversion
   and pixel mapping must match exactly.
3. Closed forms and invariants for what those leave open.
"""
import glob
import math
import os

import numpy as np
import pytest

import oracle_lib as ol

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sc_from(arr):
    R, L, ny, nz, ps, sz, sdl = arr
    return ol.SimpleScanner(R, L, int(ny), int(nz), ps, ps, sz, sdl)


# --- 1. reference unit-test goldens -------------------------------------------

@pytest.mark.parametrize("angle,y,dz,dy,ref", [
    (math.pi / 6, 0.1, 0.0, 0.0, 184.304974530866),
    (math.pi / 6, 0.1, 0.05, 0.1, 5.272096361578383e-6),
    (math.pi / 6, 0.1, -0.05, -0.05, 0.0338473429952690),
    (math.pi / 4, -0.3, 0.0, 0.0, 95.87658397191050),
    (math.pi / 4, -0.3, -0.05, -0.05, 0.1719727908599724),
])
def test_gaussian_kernel_goldens(angle, y, dz, dy, ref):
    """gaussian_kernel_test.cpp:15-51 (float kernel, epsilon 1e-6)."""
    o = ol.oracle()
    t = np.float32(math.tan(angle))
    sec = np.float32(1 / math.cos(angle))
    v32 = o.oracle_kernel_f32(0.01, 0.04, np.float32(y), t, sec, np.float32(0.35),
                              np.float32(dz), np.float32(dy))
    v64 = o.oracle_kernel_f64(0.01, 0.04, y, math.tan(angle), 1 / math.cos(angle),
                              0.35, dz, dy)
    assert v32 == pytest.approx(ref, rel=1e-6)
    assert v64 == pytest.approx(ref, rel=1e-12)


def test_gaussian_kernel_far_point_is_zero():
    o = ol.oracle()
    v = o.oracle_kernel_f64(0.01, 0.04, -0.3, 1.0, math.sqrt(2), 0.35, 0.05, 0.1)
    assert abs(v) < 1e-15


@pytest.mark.parametrize("R,L,z,y,ref", [
    (500, 1000, 0, 0, 0.5), (500, 1000, 50, 0, 0.46652458328685176),
    (500, 1000, -50, 100, 0.4410019151324715),
    (500, 1000, -450, -200, 0.07526632771111386),
    (450, 200, 0, 0, 0.1392089745461279), (450, 200, 50, 0, 0.07044657495455454),
    (450, 200, -50, 100, 0.07402517367717103),
    (450, 200, -70, -200, 0.05269621503719814),
])
def test_sensitivity_goldens(R, L, z, y, ref):
    """kernel_test.cpp:31-49 (double, 1e-13)."""
    assert ol.oracle().oracle_sensitivity_point_f64(R, L, z, y) == pytest.approx(ref, rel=1e-13)


@pytest.mark.parametrize("angle,bby,bbz", [
    (0.0470448, 94.3954, 21.6737), (-0.594145, 78.3053, 56.9959),
    (0.20029, 92.6108, 28.3458), (-0.571667, 79.4745, 55.3539),
    (-0.420542, 86.266, 44.0276),
])
def test_bbox_goldens(angle, bby, bbz):
    """kernel_test.cpp:81-113: sigma_z=10, sigma_dl=63, 1e-4."""
    out = np.zeros(6, np.float64)
    ol.oracle().oracle_ellipse_bb_f64(10.0, 63.0, math.tan(angle), ol.P(out))
    assert out[4] == pytest.approx(bby, rel=1e-4)
    assert out[5] == pytest.approx(bbz, rel=1e-4)


@pytest.mark.parametrize("y,z,deg,tan,zu,zd,dl", [
    (10.0, 20.0, 7.0, 0.1227845609029046, 74.02520679727803, -36.480898015336116,
     -20.15019650917697),
    (-10.0, 37.0, -5.0, -0.08748866352592401, -3.244785221925042, 75.49501195140655,
     20.076396750866948),
])
def test_response_goldens(y, z, deg, tan, zu, zd, dl):
    """response_test.cpp:11-55 (R=450): image->projection space and back."""
    R = 450.0
    t = math.tan(deg * math.pi / 180)
    assert t == pytest.approx(tan, rel=1e-13)
    # scanner.h:107-113 projection (the inverse of what the oracle restates)
    pzu, pzd, pdl = z + (R - y) * t, z - (R + y) * t, -2 * y * math.sqrt(t * t + 1)
    assert (pzu, pzd, pdl) == pytest.approx((zu, zd, dl), rel=1e-13)
    out = np.zeros(3, np.float64)
    ol.oracle().oracle_tan_y_z_f64(R, ol.P(np.array([zu, zd, dl])), ol.P(out))
    assert out[0] == pytest.approx(tan, rel=1e-13)
    assert out[1] == pytest.approx(y, rel=1e-13)
    assert out[2] == pytest.approx(z, rel=1e-13)


def _pixel_at(sc, z, y, f64=True):
    out = np.zeros(2, np.int32)
    s = ol.oscanner(sc)
    import ctypes as C
    fn = ol.oracle().oracle_pixel_at_f64 if f64 else ol.oracle().oracle_pixel_at_f32
    fn.argtypes = [C.POINTER(ol.OScanner), C.c_double if f64 else C.c_float,
                   C.c_double if f64 else C.c_float, C.c_void_p]
    fn(C.byref(s), z, y, ol.P(out))
    return tuple(out)


def _pixel_center(sc, x, y):
    import ctypes as C
    out = np.zeros(2, np.float64)
    fn = ol.oracle().oracle_pixel_center_f64
    fn.argtypes = [C.POINTER(ol.OScanner), C.c_int32, C.c_int32, C.c_void_p]
    fn(C.byref(ol.oscanner(sc)), x, y, ol.P(out))
    return tuple(out)


def test_pixel_at_goldens():
    """scanner_test.cpp:15-42, including truncation at the borders."""
    s = ol.SimpleScanner(500, 1000, 200, 200, 5, 5, 10, 63)
    cases = [((0., 0.), (100, 100)), ((-0.00001, 0.00001), (99, 99)),
             ((-500., 500.), (0, 0)), ((500., 500.), (200, 0)),
             ((-500., -500.), (0, 200)), ((500., -500.), (200, 200)),
             ((-499.99999, 499.99999), (0, 0)), ((499.99999, 499.99999), (199, 0)),
             ((-499.99999, -499.99999), (0, 199)), ((499.99999, -499.99999), (199, 199))]
    for (z, y), px in cases:
        assert _pixel_at(s, z, y) == px, (z, y)


def test_pixel_center_goldens():
    """scanner_test.cpp:44-66 (pixel sizes 2R/n_y and L/n_z = 1 for the
    six-argument constructor)."""
    s = ol.SimpleScanner(100, 200, 200, 200, 1.0, 1.0, 10, 63)
    assert _pixel_center(s, 100, 100) == (0.5, -0.5)
    assert _pixel_center(s, 99, 99) == (-0.5, 0.5)
    assert _pixel_center(s, 0, 0) == (-99.5, 99.5)
    for p in [(100, 100), (99, 99), (0, 0)]:
        assert _pixel_at(s, *_pixel_center(s, *p)) == p
    s2 = ol.SimpleScanner(500, 1000, 200, 200, 5, 5, 10, 63)
    for (z, y), c in [((0., 0.), (2.5, -2.5)), ((-6., 3.), (-7.5, 2.5)),
                      ((6., 3.), (7.5, 2.5)), ((-6., -3.), (-7.5, -2.5)),
                      ((6., -3.), (7.5, -2.5))]:
        assert _pixel_center(s2, *_pixel_at(s2, z, y)) == c


# --- 2. golden vectors from the compiled reference ------------------------------

FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "mlem_*.npz")))


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_mlem_bitwise_vs_reference(path):
    f = np.load(path)
    sc = sc_from(f["scanner"])
    rho = f["rho"]
    out = ol.mlem_f32(sc, f["events"], rho.shape[0], n_threads=1)
    for i in range(rho.shape[0]):
        assert np.array_equal(out[i].view(np.uint32), rho[i].view(np.uint32)), i


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_mlem_threads_and_fp64_vs_reference(path):
    f = np.load(path)
    sc = sc_from(f["scanner"])
    rho = f["rho"].astype(np.float64)
    out = ol.mlem_f32(sc, f["events"], rho.shape[0], n_threads=4)
    out64 = ol.mlem_f64(sc, f["events"], rho.shape[0], n_threads=4)
    for i in range(rho.shape[0]):
        n = np.linalg.norm(rho[i])
        assert np.linalg.norm(out[i] - rho[i]) / n < 1e-6
        assert np.linalg.norm(out64[i] - rho[i]) / n < 1e-5


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_sensitivity_vs_reference(path):
    f = np.load(path)
    sc = sc_from(f["scanner"])
    assert np.array_equal(ol.sensitivity_f32(sc, 0), f["sensitivity"])
    assert np.max(np.abs(ol.sensitivity_f32(sc, 1) - f["pixel_sensitivity"])) < 1e-7


def test_point_values_vs_reference():
    f = np.load(os.path.join(GOLDEN, "points.npz"))
    o = ol.oracle()
    for (y, t, dz, dy), (sec, ref) in zip(f["kernel_in"], f["kernel_out"]):
        v = o.oracle_kernel_f32(0.01, 0.04, np.float32(y), np.float32(t),
                                np.float32(sec), np.float32(0.43), np.float32(dz),
                                np.float32(dy))
        assert np.float32(v) == np.float32(ref)
    for t, ref in zip(f["bb_tan"], f["bb_out"]):
        out = np.zeros(6, np.float32)
        o.oracle_ellipse_bb_f32(0.01, 0.04, t, ol.P(out))
        assert np.array_equal(out, ref)
    for ev, ref in zip(f["tyz_in"], f["tyz_out"]):
        out = np.zeros(3, np.float32)
        o.oracle_tan_y_z_f32(0.43, ol.P(np.ascontiguousarray(ev)), ol.P(out))
        assert np.array_equal(out, ref)
    s = ol.SimpleScanner(0.43, 0.5, 512, 512, 0.001, 0.001, 0.01, 0.04)
    for (z, y), ref in zip(f["pixel_at_in"], f["pixel_at_out"]):
        assert _pixel_at(s, float(z), float(y), f64=False) == tuple(ref)


FOOTPRINTS = np.load(os.path.join(GOLDEN, "footprints.npz"))


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_footprint_vs_reference(path):
    """oracle_footprint, per event, against the reference's own footprints
    (tests/golden/footprints.npz, from the reference compiled from its
    headers): the ellipse pixels are those bb_update writes for the event
    alone, the box the clamped box it scans; and against the MLEM oracle's
    own counts (the path pinned bit for bit above), per event and in
    total."""
    f = np.load(path)
    sc = sc_from(f["scanner"])
    name = os.path.basename(path)[len("mlem_"):-len(".npz")]
    want_box, want_ell = FOOTPRINTS[f"box_{name}"], FOOTPRINTS[f"ell_{name}"]
    ev = f["events"][:len(want_box)]
    nb, ne = ol.footprint(sc, ev)
    assert np.array_equal(nb, want_box)
    assert np.array_equal(ne, want_ell)
    assert ne.sum() > 0 and np.all(ne <= nb)
    _, cnt = ol.mlem_f32(sc, ev, 1, counts=True)
    assert (int(cnt[0]), int(cnt[1])) == (int(nb.sum()), int(ne.sum()))
    for k in range(0, len(ev), 41):
        _, c1 = ol.mlem_f32(sc, ev[k:k + 1], 1, counts=True)
        assert (int(c1[0]), int(c1[1])) == (int(nb[k]), int(ne[k])), k


def test_footprint_brute_force_tiny():
    """A tiny image where every box pixel is tested by hand in FP64 with
    the reference's rules: the count of pixels with q = A dy^2 + B dy dz +
    C dz^2 <= 9 matches oracle_footprint wherever no pixel's q lies within
    1e-5 of the threshold (there the float test may round either way)."""
    rng = np.random.default_rng(5)
    sc = ol.SimpleScanner(0.43, 0.5, 12, 12, 0.02, 0.02, 0.01, 0.04)
    o = ol.oracle()
    n_checked = 0
    for _ in range(200):
        ev = np.array([rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1),
                       rng.normal(0, 0.08)], np.float32)
        tyz = np.zeros(3, np.float64)
        o.oracle_tan_y_z_f64(0.43, ol.P(ev.astype(np.float64)), ol.P(tyz))
        t, y, z = tyz
        bb = np.zeros(6, np.float64)
        o.oracle_ellipse_bb_f64(0.01, 0.04, t, ol.P(bb))
        _, A, B, C, bby, bbz = bb
        cx = int(np.trunc((z + 0.12) / 0.02))
        cy = int(np.trunc((0.12 - y) / 0.02))
        hw, hh = int(bbz / 0.02 + 0.5), int(bby / 0.02 + 0.5)
        x0, x1 = min(11, max(0, cx - hw)), min(11, max(0, cx + hw))
        y0, y1 = min(11, max(0, cy - hh)), min(11, max(0, cy + hh))
        qs = [A * (0.12 - 0.01 - iy * 0.02 - y) ** 2 +
              B * (0.12 - 0.01 - iy * 0.02 - y) * (-0.12 + 0.01 + iz * 0.02 - z) +
              C * (-0.12 + 0.01 + iz * 0.02 - z) ** 2
              for iy in range(y0, y1) for iz in range(x0, x1)]
        if any(abs(q - 9) < 1e-5 for q in qs):
            continue
        nb, ne = ol.footprint(sc, ev[None, :])
        assert nb[0] == len(qs) and ne[0] == sum(q <= 9 for q in qs)
        n_checked += 1
    assert n_checked > 150


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_pixel_center_f32_vs_reference(path):
    """oracle_pixel_center_f32 for every pixel against the reference's
    Scanner::pixel_center in float (tests/golden/footprints.npz)."""
    import ctypes as C
    f = np.load(path)
    sc = sc_from(f["scanner"])
    name = os.path.basename(path)[len("mlem_"):-len(".npz")]
    want = FOOTPRINTS[f"centre_{name}"]
    fn = ol.oracle().oracle_pixel_center_f32
    s = ol.oscanner(sc)
    got = np.zeros_like(want)
    for y in range(sc.n_y_pixels):
        for x in range(sc.n_z_pixels):
            o = np.zeros(2, np.float32)
            fn(C.byref(s), x, y, ol.P(o))
            got[y, x] = o
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.skipif(not ol.have_ref(), reason="reference harness not built here")
def test_random_cases_vs_live_reference():
    """Fresh random cases against the compiled reference (only where the
    reference tree was present to build oracle/_ref)."""
    rng = np.random.default_rng(2024)
    for n_pix, ps in [(37, 0.013), (80, 0.006)]:
        ev = np.stack([rng.uniform(-0.3, 0.3, 3000), rng.uniform(-0.3, 0.3, 3000),
                       rng.normal(0, 0.2, 3000)], 1).astype(np.float32)
        sc = ol.SimpleScanner(0.43, 0.5, n_pix, n_pix, ps, ps, 0.01, 0.04)
        a = ol.mlem_f32(sc, ev, 2, n_threads=1)
        b = ol.ref_mlem(sc, ev, 2, n_threads=1)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


# --- 3. closed forms and invariants ---------------------------------------------

def test_sensitivity_closed_forms():
    o = ol.oracle()
    R, L = 0.43, 0.5
    assert o.oracle_sensitivity_point_f64(R, L, 0, 0) == pytest.approx(
        2 / math.pi * math.atan(L / (2 * R)), rel=1e-14)
    for z, y in [(0.1, 0.05), (-0.2, 0.3), (0.03, -0.1)]:
        s = o.oracle_sensitivity_point_f64(R, L, z, y)
        assert s == pytest.approx(o.oracle_sensitivity_point_f64(R, L, -z, y), rel=1e-14)
        assert s == pytest.approx(o.oracle_sensitivity_point_f64(R, L, z, -y), rel=1e-14)


def test_bbox_special_case_tan0():
    """At tan = 0: bb_y = 1.5 sigma_dl, bb_z = 3 sigma_z / sqrt(2)."""
    out = np.zeros(6, np.float64)
    ol.oracle().oracle_ellipse_bb_f64(0.01, 0.04, 0.0, ol.P(out))
    assert out[4] == pytest.approx(1.5 * 0.04, rel=1e-14)
    assert out[5] == pytest.approx(3 * 0.01 / math.sqrt(2), rel=1e-14)


def _brute_single_event(sc, ev, rho):
    """Independent FP64 statement of one event's update from the reference's
    formulas: support by brute force over all pixels with the box rule, kernel
    from the covariance form K = c/sqrt(a^T S a) exp(-b^T S b / 2)."""
    zu, zd, dl = map(float, ev)
    R = sc.radius
    t = (zu - zd) / (2 * R)
    y = -0.5 * dl / math.sqrt(1 + t * t)
    z = 0.5 * (zu + zd + 2 * y * t)
    sec = math.sqrt(1 + t * t)
    isz, isdl = 1 / sc.sigma_z ** 2, 1 / sc.sigma_dl ** 2
    Sinv = np.diag([isz, isz, isdl])
    c = 1 / (2 * math.pi ** 2) * math.sqrt(isz * isz * isdl)
    A = 4 * sec * sec * isdl + 2 * t * t * isz
    B = -4 * t * isz
    C = 2 * isz
    bb_y = 3 / math.sqrt(A - (B / 2) ** 2 / C)
    bb_z = 3 / math.sqrt(C - (B / 2) ** 2 / A)
    W, H, pw, ph = sc.n_z_pixels, sc.n_y_pixels, sc.pixel_width, sc.pixel_height
    tl_z, tl_y = -W * pw / 2, H * ph / 2
    cx, cy = int((z - tl_z) / pw), int((tl_y - y) / ph)
    hw, hh = int(bb_z / pw + 0.5), int(bb_y / ph + 0.5)
    x0, x1 = [min(max(v, 0), W - 1) for v in (cx - hw, cx + hw)]
    y0, y1 = [min(max(v, 0), H - 1) for v in (cy - hh, cy + hh)]
    out = np.zeros((H, W))
    num = {}
    for iy in range(y0, y1):
        for iz in range(x0, x1):
            pz = tl_z + pw / 2 + iz * pw
            py = tl_y - ph / 2 - iy * ph
            dz, dy = pz - z, py - y
            if A * dy * dy + B * dy * dz + C * dz * dz > 9:
                continue
            yp = py
            a = np.array([-(yp - R) * sec * sec, -(yp + R) * sec * sec, -2 * yp * sec * t])
            b = np.array([dz - dy * t, dz - dy * t, -2 * dy * sec])
            K = c / math.sqrt(a @ Sinv @ a) * math.exp(-0.5 * (b @ Sinv @ b))
            num[(iy, iz)] = K * rho[iy, iz]
    d = sum(num.values())
    for (iy, iz), v in num.items():
        out[iy, iz] = v / d
    return out


def test_single_event_closed_form():
    sc = ol.SimpleScanner(0.43, 0.5, 40, 40, 0.01, 0.01, 0.01, 0.04)
    rng = np.random.default_rng(1)
    rho = rng.uniform(0.5, 2.0, (40, 40))
    for ev in [(0.02, -0.03, 0.1), (-0.1, 0.05, -0.2), (0.0, 0.0, 0.0)]:
        ev = np.array([ev], np.float32)
        want = _brute_single_event(sc, ev[0], rho)
        got = ol.mlem_f64(sc, ev, 1, rho0=rho)[0]
        assert np.allclose(got, want, rtol=1e-9, atol=1e-12)
        assert got.sum() == pytest.approx(1.0, abs=1e-12)


def test_mlem_invariants():
    f = np.load(FIXTURES[0])
    sc = sc_from(f["scanner"])
    ev = f["events"]
    it = ol.mlem_f32(sc, ev, 2, n_threads=4)
    # every event with a positive denominator distributes exactly 1
    lam = ol.denominators_f32(sc, ev, np.ones((sc.n_y_pixels, sc.n_z_pixels), np.float32))
    assert it[0].sum(dtype=np.float64) == pytest.approx(np.count_nonzero(lam > 0), rel=1e-5)
    assert np.all(it >= 0)
    assert np.all(it[:, -1, :] == 0) and np.all(it[:, :, -1] == 0)
    # event order does not matter beyond rounding
    perm = np.random.default_rng(0).permutation(len(ev))
    it2 = ol.mlem_f32(sc, ev[perm], 2, n_threads=4)
    assert np.linalg.norm(it2[1] - it[1]) / np.linalg.norm(it[1]) < 1e-6


def test_planted_mistake_is_caught():
    """A plausible slip (swapped z_u/z_d, i.e. a sign error in tan) must be
    visible against the reference fixture."""
    f = np.load(FIXTURES[0])
    sc = sc_from(f["scanner"])
    ev = f["events"][:, [1, 0, 2]]
    out = ol.mlem_f32(sc, ev, 1, n_threads=1)[0]
    rho = f["rho"][0]
    assert np.linalg.norm(out - rho) / np.linalg.norm(rho) > 1e-2


def test_loglik_gradient_matches_finite_differences():
    """FP64 log-likelihood gradient of the oracle against central differences
    (frozen support), image and both sigmas."""
    f = np.load(FIXTURES[0])
    sc = sc_from(f["scanner"])
    ev = f["events"][:2000]
    rng = np.random.default_rng(3)
    rho = rng.uniform(0.5, 1.5, (sc.n_y_pixels, sc.n_z_pixels))
    ll, _, g, gs = ol.loglik_f64(sc, ev, rho)
    # image: 8 random pixels
    idx = rng.choice(rho.size, 8, replace=False)
    for i in idx:
        h = 1e-4 * rho.flat[i]
        rp, rm = rho.copy(), rho.copy()
        rp.flat[i] += h
        rm.flat[i] -= h
        fd = (ol.loglik_f64(sc, ev, rp, want_grad=False)[0] -
              ol.loglik_f64(sc, ev, rm, want_grad=False)[0]) / (2 * h)
        assert g.flat[i] == pytest.approx(fd, rel=1e-5, abs=1e-7)
    # sigmas, support frozen at the nominal sigma
    s0 = (sc.sigma_z, sc.sigma_dl)
    for k in range(2):
        h = 1e-4 * s0[k]
        sp, sm = list(s0), list(s0)
        sp[k] += h
        sm[k] -= h
        fp = ol.loglik_f64(sc, ev, rho, sk=tuple(sp), ss=s0, want_grad=False)[0]
        fm = ol.loglik_f64(sc, ev, rho, sk=tuple(sm), ss=s0, want_grad=False)[0]
        assert gs[k] == pytest.approx((fp - fm) / (2 * h), rel=1e-5)


def test_loglik_gradient_is_mlem_backprojection():
    """grad_rho l + 1 = rho'/rho (MLEM without sensitivity)."""
    f = np.load(FIXTURES[0])
    sc = sc_from(f["scanner"])
    ev = f["events"]
    rho = f["rho"][0].astype(np.float64)
    _, _, g, _ = ol.loglik_f64(sc, ev, rho)
    nxt = ol.mlem_f64(sc, ev, 1, rho0=rho)[0]
    m = rho > 1e-3
    assert np.allclose((g + 1)[m] * rho[m], nxt[m], rtol=1e-9, atol=1e-9)
