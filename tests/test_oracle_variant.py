"""The oracle's restatement of the reference --gpu iteration: the analytic
kernel (analytic_kernel.h:28-62) and the MLEM update with the sensitivity in
the denominator (cuda/reconstruction.cu:59-113, warp_granularity.cuh:42-153).

Pinned by the compiled reference's AnalyticKernel (oracle/_ref, where the
reference tree was present to build it), by closed forms (zero distance, and
tan = 0 on the row through the centre at y = 0, where the analytic kernel is
the Gaussian one), by an independent numpy statement in covariance form for
single events, and by the update's invariant sum_p s_p rho'_p = number of
events with a positive denominator.  The reference's own unit tests hold no
analytic-kernel goldens (kernel_test.cpp:69-77 are compiled out)."""
import math

import numpy as np
import pytest

import oracle_lib as ol

SZ, SDL, R = 0.01, 0.04, 0.43


def _cases(n, seed):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        t = rng.uniform(-0.6, 0.6)
        y = rng.uniform(-0.2, 0.2)
        dz = rng.uniform(-0.03, 0.03)
        dy = rng.uniform(-0.08, 0.08)
        yield y, t, math.sqrt(1 + t * t), dz, dy


@pytest.mark.skipif(not ol.have_ref(), reason="reference harness not built here")
def test_analytic_kernel_vs_compiled_reference():
    o, r = ol.oracle(), ol.ref()
    for y, t, sec, dz, dy in _cases(400, 5):
        want = r.ref_strip_analytic_kernel_f64(SZ, SDL, y, t, sec, R, dz, dy)
        got = o.oracle_analytic_kernel_f64(SZ, SDL, y, t, sec, R, dz, dy)
        assert got == pytest.approx(want, rel=1e-13, abs=1e-300)
        w32 = r.ref_strip_analytic_kernel_f32(SZ, SDL, y, t, sec, R, dz, dy)
        g32 = o.oracle_analytic_kernel_f32(SZ, SDL, y, t, sec, R, dz, dy)
        assert g32 == pytest.approx(w32, rel=2e-4, abs=1e-30 * max(w32, 1.0))


def test_analytic_kernel_closed_forms():
    o = ol.oracle()
    for y, t, sec, _, _ in _cases(50, 6):
        # zero distance: b = 0, so K = c / sqrt(a S a) as for the Gaussian
        a = o.oracle_analytic_kernel_f64(SZ, SDL, y, t, sec, R, 0.0, 0.0)
        g = o.oracle_kernel_f64(SZ, SDL, y, t, sec, R, 0.0, 0.0)
        assert a == pytest.approx(g, rel=1e-14)
    for dz in (-0.02, -0.005, 0.0, 0.011, 0.025):
        # tan = 0, y = 0, dy = 0: o = 0 and b.S.a = 0, the Gaussian again
        a = o.oracle_analytic_kernel_f64(SZ, SDL, 0.0, 0.0, 1.0, R, dz, 0.0)
        g = o.oracle_kernel_f64(SZ, SDL, 0.0, 0.0, 1.0, R, dz, 0.0)
        assert a == pytest.approx(g, rel=1e-13)
    # elsewhere the o and (b.S.a)^2/norm terms make a difference
    diffs = [abs(o.oracle_analytic_kernel_f64(SZ, SDL, y, t, sec, R, dz, dy) /
                 o.oracle_kernel_f64(SZ, SDL, y, t, sec, R, dz, dy) - 1)
             for y, t, sec, dz, dy in _cases(50, 7)]
    assert max(diffs) > 1e-3


def _brute_single_event(sc, ev, rho, sens):
    """rho'_p = K_p rho_p / sum_q K_q rho_q s_q over the box and 3-sigma
    ellipse, K the analytic kernel in covariance form (numpy matrices)."""
    zu, zd, dl = map(float, ev)
    t = (zu - zd) / (2 * R)
    y = -0.5 * dl / math.sqrt(1 + t * t)
    z = 0.5 * (zu + zd + 2 * y * t)
    sec = math.sqrt(1 + t * t)
    isz, isdl = 1 / SZ ** 2, 1 / SDL ** 2
    S = np.diag([isz, isz, isdl])
    c = 1 / (2 * math.pi ** 2) * math.sqrt(isz * isz * isdl)
    A, B, Cc = 4 * sec * sec * isdl + 2 * t * t * isz, -4 * t * isz, 2 * isz
    bb_y = 3 / math.sqrt(A - (B / 2) ** 2 / Cc)
    bb_z = 3 / math.sqrt(Cc - (B / 2) ** 2 / A)
    W, H, pw, ph = sc.n_z_pixels, sc.n_y_pixels, sc.pixel_width, sc.pixel_height
    tl_z, tl_y = -W * pw / 2, H * ph / 2
    cx, cy = int((z - tl_z) / pw), int((tl_y - y) / ph)
    hw, hh = int(bb_z / pw + 0.5), int(bb_y / ph + 0.5)
    x0, x1 = [min(max(v, 0), W - 1) for v in (cx - hw, cx + hw)]
    y0, y1 = [min(max(v, 0), H - 1) for v in (cy - hh, cy + hh)]
    num, den = {}, 0.0
    for iy in range(y0, y1):
        for iz in range(x0, x1):
            dz = tl_z + pw / 2 + iz * pw - z
            dy = tl_y - ph / 2 - iy * ph - y
            if A * dy * dy + B * dy * dz + Cc * dz * dz > 9:
                continue
            yp = dy + y
            a = np.array([-(yp - R) * sec * sec, -(yp + R) * sec * sec, -2 * yp * sec * t])
            o = np.array([a[0] * t, a[1] * t, -yp * sec * (1 + 2 * t * t)])
            b = np.array([dz - dy * t, dz - dy * t, -2 * dy * sec])
            norm = a @ S @ a + 2 * (o @ S @ b)
            K = c / math.sqrt(norm) * math.exp(-0.5 * (b @ S @ b - (b @ S @ a) ** 2 / norm))
            num[(iy, iz)] = K * rho[iy, iz]
            den += K * rho[iy, iz] * sens[iy, iz]
    out = np.zeros((H, W))
    for (iy, iz), v in num.items():
        out[iy, iz] = v / den
    return out


def test_variant_single_event_closed_form():
    sc = ol.SimpleScanner(R, 0.5, 40, 40, 0.01, 0.01, SZ, SDL)
    rng = np.random.default_rng(3)
    rho = rng.uniform(0.5, 2.0, (40, 40))
    sens = rng.uniform(0.2, 0.9, (40, 40))
    for ev in [(0.02, -0.03, 0.1), (-0.1, 0.05, -0.2), (0.0, 0.0, 0.0), (0.15, -0.2, 0.05)]:
        ev = np.array([ev], np.float32)
        want = _brute_single_event(sc, ev[0], rho, sens)
        got = ol.mlem_variant_f64(sc, ev, 1, kernel=1, sens=sens, rho0=rho)[0]
        assert np.allclose(got, want, rtol=1e-9, atol=1e-12)
        assert (got * sens).sum() == pytest.approx(1.0, abs=1e-12)


def test_variant_reduces_to_cpu_update():
    """Gaussian kernel and no sensitivity: the CPU path's FP64 twin, bit for
    bit."""
    sc = ol.SimpleScanner(R, 0.5, 48, 48, 0.008, 0.008, SZ, SDL)
    rng = np.random.default_rng(4)
    ev = np.stack([rng.uniform(-0.15, 0.15, 2000), rng.uniform(-0.15, 0.15, 2000),
                   rng.normal(0, 0.1, 2000)], 1).astype(np.float32)
    a = ol.mlem_variant_f64(sc, ev, 2, kernel=0, sens=None, n_threads=1)
    b = ol.mlem_f64(sc, ev, 2, n_threads=1)
    assert np.array_equal(a, b)


def test_variant_invariant():
    sc = ol.SimpleScanner(R, 0.5, 48, 48, 0.008, 0.008, SZ, SDL)
    rng = np.random.default_rng(5)
    ev = np.stack([rng.uniform(-0.15, 0.15, 3000), rng.uniform(-0.15, 0.15, 3000),
                   rng.normal(0, 0.1, 3000)], 1).astype(np.float32)
    sens = ol.sensitivity_f32(sc, 1).astype(np.float64)
    it = ol.mlem_variant_f64(sc, ev, 1, kernel=1, sens=sens, n_threads=4)[0]
    n_pos = np.count_nonzero(ol.denominators_f32(
        sc, ev, np.ones((48, 48), np.float32)) > 0)
    assert (it * sens).sum() == pytest.approx(n_pos, rel=1e-9)
