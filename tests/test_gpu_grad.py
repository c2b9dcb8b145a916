"""GPU forward projector, adjoint and sigma gradient against the FP64 oracle.

The oracle's gradient is itself checked against FP64 central finite
differences on the CPU (test_oracle_reference.py); here the GPU gradient is
compared with it, the adjoint is checked by the dot-product identity
<A^T g, rho> = <g, A rho>, and the image gradient by the MLEM identity
grad + 1 = rho'/rho (reconstruction.h:220-298 backprojection).
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def setup(pkg, n_pix=64, size=0.008, n=20000, seed=3, path=None):
    sc = pkg.Scanner(0.43, 0.5, n_pix, n_pix, size, size, 0.01, 0.04)
    ev = pkg.phantom_events(sc, pkg.workloads.PHANTOM_3, n, seed=seed)
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.5, 1.5, (n_pix, n_pix)).astype(np.float32)
    kw = {} if path is None else {"path": path}
    return sc, ev, rho, kw


@pytest.mark.parametrize("path", ["auto", "per_event"])
def test_gradient_vs_oracle(gpu_pkg, path):
    pkg = gpu_pkg
    p = pkg.PATH_AUTO if path == "auto" else pkg.PATH_PER_EVENT
    sc, ev, rho, _ = setup(pkg)
    with pkg.Reconstruction(sc, path=p) as rec:
        rec.set_events(ev)
        rec.rho = rho
        lam = rec.forward()
        grad, gsig, ll = rec.backward(want_loglik=True)
    ll_ref, lam_ref, g_ref, gs_ref = ol.loglik_f64(sc, ev, rho, want_lambda=True)
    print(f"case {path}: lambda {rel(lam, lam_ref):.2e} grad {rel(grad, g_ref):.2e} "
          f"sigma {gsig} vs {gs_ref} loglik {ll} vs {ll_ref}", flush=True)
    assert rel(lam, lam_ref) < 2e-6
    assert rel(grad, g_ref) < 1e-5
    assert gsig[0] == pytest.approx(gs_ref[0], rel=2e-4)
    assert gsig[1] == pytest.approx(gs_ref[1], rel=2e-4)
    assert ll == pytest.approx(ll_ref, rel=1e-6)


@pytest.mark.parametrize("scale", [1.0, 1e-6, 1e-12])
def test_gradient_scaled_image(gpu_pkg, scale):
    """Log-likelihood gradient for images scaled down so that K / lambda
    grows by 1/scale: the single-conversion, split-conversion and direct
    global paths of the fused gradient against the FP64 oracle."""
    pkg = gpu_pkg
    sc, ev, rho, _ = setup(pkg, n=10000)
    rho = (rho * scale).astype(np.float32)
    with pkg.Reconstruction(sc) as rec:
        rec.set_events(ev)
        rec.rho = rho
        grad, gsig, _ = rec.backward()
    _, _, g_ref, gs_ref = ol.loglik_f64(sc, ev, rho)
    print(f"case scale {scale}: grad {rel(grad, g_ref):.2e} sigma {gsig} vs {gs_ref}",
          flush=True)
    assert rel(grad, g_ref) < 1e-5
    assert gsig[0] == pytest.approx(gs_ref[0], rel=2e-4)
    assert gsig[1] == pytest.approx(gs_ref[1], rel=2e-4)


def test_adjoint_dot_product(gpu_pkg):
    pkg = gpu_pkg
    sc, ev, rho, _ = setup(pkg, n=15000)
    g = np.random.default_rng(9).normal(size=len(ev)).astype(np.float32)
    with pkg.Reconstruction(sc) as rec:
        rec.set_events(ev)
        rec.rho = rho
        lam = rec.forward()
        atg, gsig, _ = rec.backward(g=g)
        # linearity in g
        atg2, _, _ = rec.backward(g=2 * g)
    lhs = float(np.dot(atg.astype(np.float64).ravel(), rho.astype(np.float64).ravel()))
    rhs = float(np.dot(g.astype(np.float64), lam.astype(np.float64)))
    print(f"<A^T g, rho> = {lhs:.9g}  <g, A rho> = {rhs:.9g}", flush=True)
    assert lhs == pytest.approx(rhs, rel=1e-5)
    assert np.allclose(atg2, 2 * atg, rtol=1e-5, atol=1e-6)


def test_gradient_is_mlem_backprojection(gpu_pkg):
    """grad l + 1 = rho'/rho where rho' is one MLEM iteration from rho."""
    pkg = gpu_pkg
    sc, ev, rho, _ = setup(pkg)
    with pkg.Reconstruction(sc) as rec:
        rec.set_events(ev)
        rec.rho = rho
        grad, _, _ = rec.backward(want_sigma=False)
        rec(1)
        nxt = rec.rho
    m = rho > 0
    assert rel((grad + 1)[m] * rho[m], nxt[m]) < 1e-5


def test_frozen_support_sigma_fd(gpu_pkg):
    """GPU sigma gradient against FP64 central differences of the
    log-likelihood in sigma with the support frozen at the unperturbed
    sigma (the FP64 restatement evaluated at sigma +- h), and the GPU
    log-likelihood at perturbed sigma against the FP64 value there.

    Differencing the FP32 GPU log-likelihood itself is not a usable check:
    the sigma-dependent constants of the kernel (normalisation, 1/sigma^2)
    are rounded to FP32, which makes the FP32 log-likelihood a step function
    of sigma at the 1e-7 relative level, i.e. O(1e-3) noise in a difference
    quotient with h = 1e-3 sigma."""
    pkg = gpu_pkg
    sc, ev, rho, _ = setup(pkg, n=20000)
    s0 = (0.01, 0.04)
    with pkg.Reconstruction(sc) as rec:
        rec.set_events(ev)
        rec.rho = rho
        _, gsig, _ = rec.backward(want_rho=False)
        fd = []
        for k in range(2):
            h = 1e-4 * s0[k]
            vals = []
            for sgn in (+1, -1):
                sk = list(s0)
                sk[k] += sgn * h
                ref = ol.loglik_f64(sc, ev, rho, sk=tuple(sk), ss=s0,
                                    want_grad=False)[0]
                rec.set_sigma(sk[0], sk[1], freeze_support=True)
                got = rec.backward(want_rho=False, want_sigma=False,
                                   want_loglik=True)[2]
                print(f"sigma {sk}: loglik {got} FP64 {ref}", flush=True)
                assert got == pytest.approx(ref, rel=1e-6)
                vals.append(ref)
            fd.append((vals[0] - vals[1]) / (2 * h))
        rec.set_sigma(s0[0], s0[1], freeze_support=True)
        _, gsig2, _ = rec.backward(want_rho=False)
    print(f"sigma grad {gsig} FP64 fd {fd}", flush=True)
    assert gsig[0] == pytest.approx(fd[0], rel=2e-4)
    assert gsig[1] == pytest.approx(fd[1], rel=2e-4)
    # restoring sigma restores the gradient
    assert gsig2[0] == pytest.approx(gsig[0], rel=1e-6)
    assert gsig2[1] == pytest.approx(gsig[1], rel=1e-6)


def test_workload_256px_gradient(gpu_pkg):
    """256x256 pixels of 2 mm, 5e5 events of the three-ellipse phantom:
    image and sigma gradient against the FP64 oracle."""
    pkg = gpu_pkg
    sc = pkg.workload_scanner("strip-256px-2mm-5e6ev")
    ev = pkg.workload_events("strip-256px-2mm-5e6ev", n=500000)
    rho = np.ones((256, 256), np.float32)
    with pkg.Reconstruction(sc) as rec:
        rec.set_events(ev)
        rec(2)
        rho = rec.rho
        grad, gsig, ll = rec.backward(want_loglik=True)
    ll_ref, _, g_ref, gs_ref = ol.loglik_f64(sc, ev, rho, n_threads=16)
    m = rho > 1e-3 * rho.max()
    print(f"grad {rel(grad[m], g_ref[m]):.2e} sigma {gsig} {gs_ref}", flush=True)
    assert rel(grad[m], g_ref[m]) < 1e-5
    # per-event derivatives are O(1/sigma); their sum may cancel, so the
    # tolerance is scaled by N/sigma rather than by the result
    n = len(ev)
    assert gsig[0] == pytest.approx(gs_ref[0], rel=1e-3, abs=1e-5 * n / 0.01)
    assert gsig[1] == pytest.approx(gs_ref[1], rel=1e-3, abs=1e-5 * n / 0.04)
    assert ll == pytest.approx(ll_ref, rel=1e-6)


@pytest.mark.parametrize("bands", [0, 1])
def test_sigma_gradient_reproducible(gpu_pkg, bands):
    """The sigma gradient and the log-likelihood terms are summed per
    workgroup (per block of the band combine) and then in a fixed order, so
    two calls give the same bits; the fused form and the banded form agree
    with each other and with the oracle."""
    pkg = gpu_pkg
    sc, ev, rho, _ = setup(pkg, n=50000, seed=5)
    with pkg.Reconstruction(sc) as rec:
        rec.set_options(force_bands=bands)
        rec.set_events(ev)
        rec.rho = rho
        out = [rec.backward(want_loglik=True) for _ in range(3)]
        fw = [rec.backward(g=np.full(len(ev), 0.5, np.float32), want_rho=False)
              for _ in range(2)]
    ll_ref, _, g_ref, gs_ref = ol.loglik_f64(sc, ev, rho)
    print(f"bands {bands}: sigma {out[0][1]} vs {gs_ref}, loglik {out[0][2]} vs {ll_ref}",
          flush=True)
    for grad, gsig, ll in out[1:]:
        assert gsig == out[0][1]
        assert ll == out[0][2]
        assert np.array_equal(grad, out[0][0])
    assert fw[1][1] == fw[0][1]
    assert out[0][1][0] == pytest.approx(gs_ref[0], rel=2e-4)
    assert out[0][1][1] == pytest.approx(gs_ref[1], rel=2e-4)
    assert out[0][2] == pytest.approx(ll_ref, rel=1e-6)


class _EventEllipses:
    """FP64 centres and ellipse coefficients of every event (response.h:22-37,
    gaussian_kernel.h:108-137), to find the events whose ellipse may hold a
    pixel."""

    def __init__(self, sc, ev):
        e = np.asarray(ev, np.float64)
        R, sz, sdl = sc.radius, sc.sigma_z, sc.sigma_dl
        t = (e[:, 0] - e[:, 1]) / (2 * R)
        self.y = -0.5 * e[:, 2] / np.sqrt(1 + t * t)
        self.z = 0.5 * (e[:, 0] + e[:, 1] + 2 * self.y * t)
        self.A = 4 * (1 + t * t) / sdl ** 2 + 2 * t * t / sz ** 2
        self.B = -4 * t / sz ** 2
        self.C = 2 / sz ** 2
        self.sc = sc

    def near(self, iy, iz):
        """A superset (the ellipse test with a 1e-3 margin) of the events
        whose ellipse holds pixel (iy, iz); the others have a lambda that
        does not depend on the pixel."""
        sc = self.sc
        pw, ph = sc.pixel_width, sc.pixel_height
        zp = sc.center_z - sc.n_z_pixels * pw / 2 + pw / 2 + iz * pw
        yp = sc.center_y + sc.n_y_pixels * ph / 2 - ph / 2 - iy * ph
        dy, dz = yp - self.y, zp - self.z
        q = self.A * dy * dy + self.B * dy * dz + self.C * dz * dz
        return np.nonzero(q <= 9 * (1 + 1e-3))[0]


def test_workload_256px_fd_64_pixels(gpu_pkg):
    """256x256 pixels of 2 mm, 5e6 events of the three-ellipse phantom, the
    image after two MLEM iterations: the GPU image gradient on 64 seeded
    random pixels against FP64 central differences of the log-likelihood
    (step 1e-4 of the pixel value; only the events whose box may hold the
    pixel change, so the difference is taken over them), the GPU sigma
    gradient against FP64 central differences in sigma (step 1e-4 relative,
    support frozen), and the MLEM identity (grad + 1) rho = rho' over the
    whole image."""
    pkg = gpu_pkg
    name = "strip-256px-2mm-5e6ev"
    sc = pkg.workload_scanner(name)
    ev = pkg.workload_events(name)
    print(f"case: {name}, {len(ev)} events", flush=True)
    s0 = (sc.sigma_z, sc.sigma_dl)
    with pkg.Reconstruction(sc) as rec:
        rec.set_events(ev)
        rec(2)
        rho = rec.rho.copy()
        grad, gsig, ll = rec.backward(want_loglik=True)
        rec(1)
        nxt = rec.rho.copy()
    m = rho > 0
    assert rel((grad + 1)[m] * rho[m], nxt[m]) < 1e-5
    rng = np.random.default_rng(4)
    cand = np.argwhere(rho > 1e-2 * rho.max())
    pix = cand[rng.choice(len(cand), 64, replace=False)]
    r64 = rho.astype(np.float64)
    ells = _EventEllipses(sc, ev)
    worst = 0.0
    for iy, iz in pix:
        sub = np.ascontiguousarray(ev[ells.near(iy, iz)])
        h = 1e-4 * r64[iy, iz]
        vals = []
        for sgn in (+1, -1):
            rp = r64.copy()
            rp[iy, iz] += sgn * h
            vals.append(ol.loglik_f64(sc, sub, rp, want_grad=False, n_threads=16)[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        err = abs(grad[iy, iz] - fd) / max(abs(fd), 1.0)
        worst = max(worst, err)
        assert err < 1e-4, (iy, iz, grad[iy, iz], fd)
    print(f"64 pixels: worst |grad - fd| / max(|fd|, 1) = {worst:.2e}", flush=True)
    fd_sig = []
    for k in range(2):
        h = 1e-4 * s0[k]
        vals = []
        for sgn in (+1, -1):
            sk = list(s0)
            sk[k] += sgn * h
            vals.append(ol.loglik_f64(sc, ev, rho, sk=tuple(sk), ss=s0, want_grad=False,
                                      n_threads=16)[0])
        fd_sig.append((vals[0] - vals[1]) / (2 * h))
    print(f"sigma grad {gsig} FP64 fd {fd_sig}", flush=True)
    n = len(ev)
    assert gsig[0] == pytest.approx(fd_sig[0], rel=1e-3, abs=1e-5 * n / s0[0])
    assert gsig[1] == pytest.approx(fd_sig[1], rel=1e-3, abs=1e-5 * n / s0[1])


def test_loglik_independent_of_sigma_history(gpu_pkg):
    """The log-likelihood and sigma gradient at a sigma do not depend on the
    sigmas set before it: sigma_z values a few ulps apart share the kernel
    tables' u step (du = pixel width x s_pix rounds to the same float) while
    their s_pix differ, and each must take its own s_pix."""
    pkg = gpu_pkg
    sc, ev, rho, _ = setup(pkg)
    sz, sdl = 0.0111, 0.043
    with pkg.Reconstruction(sc) as rec:
        rec.set_events(ev)
        rec.rho = rho
        rec.set_sigma(sz, sdl, freeze_support=True)
        _, g0, ll0 = rec.backward(want_rho=False, want_loglik=True)
        print(f"case: sigma ({sz}, {sdl}), loglik {ll0!r}", flush=True)
        for k in range(1, 25):
            rec.set_sigma(sz * (1.0 + 3e-8 * k), sdl, freeze_support=True)
            rec.backward(want_rho=False, want_loglik=True)
            rec.set_sigma(sz, sdl, freeze_support=True)
            _, g, ll = rec.backward(want_rho=False, want_loglik=True)
            assert ll == ll0 and tuple(g) == tuple(g0), k


@pytest.mark.parametrize("bands", [0, 1])
def test_gradient_cold_region(gpu_pkg, bands):
    """An image with a square 1e12 times colder than the rest: the events
    whose ellipses lie in it have tiny denominators, and the values
    K / lambda_e of their rows exceed the range of the 64-bit windows, so
    those rows add to the image directly.  Image gradient, sigma gradient
    and log-likelihood against the FP64 oracle."""
    pkg = gpu_pkg
    sc, ev, rho, _ = setup(pkg, n=30000, seed=7)
    rho[20:44, 20:44] = 1e-12
    with pkg.Reconstruction(sc, force_bands=bands) as rec:
        rec.set_events(ev)
        rec.rho = rho
        grad, gsig, ll = rec.backward(want_loglik=True)
    ll_ref, _, g_ref, gs_ref = ol.loglik_f64(sc, ev, rho)
    print(f"case: bands {bands}: sigma {gsig} vs {gs_ref}, loglik {ll} vs {ll_ref}",
          flush=True)
    # the cold square holds the largest gradient values
    assert g_ref[20:44, 20:44].max() > 1e9
    assert rel(grad, g_ref) < 1e-5
    assert gsig[0] == pytest.approx(gs_ref[0], rel=2e-4)
    assert gsig[1] == pytest.approx(gs_ref[1], rel=2e-4)
    assert ll == pytest.approx(ll_ref, rel=1e-6)
