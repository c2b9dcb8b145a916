"""The reference --gpu iteration as an option: the analytic kernel
(analytic_kernel.h:28-62) and the 5-point pixel sensitivity in the
denominator (cuda/reconstruction.cu:59-113, warp_granularity.cuh:42-153),
against the oracle's FP64 restatement of the same iteration (itself pinned by
the compiled reference's AnalyticKernel and closed forms,
test_oracle_variant.py), on the fused walks (the FP32 iteration: rho times
the sensitivity in pass 1, the analytic kernel per pixel) and on the
per-event kernels.  The reference GPU build's float row division that
misplaces pixels (warp_granularity.cuh:171) is not reproduced; the
reference's unguarded 1/denominator is guarded."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def small(pkg, n_pix=64, size=0.008, n=20000, seed=3):
    sc = pkg.Scanner(0.43, 0.5, n_pix, n_pix, size, size, 0.01, 0.04)
    ev = pkg.phantom_events(sc, pkg.workloads.PHANTOM_3, n, seed=seed)
    return sc, ev


@pytest.mark.parametrize("path", ["auto", "per_event"])
@pytest.mark.parametrize("kernel,sens", [(1, True), (0, True), (1, False)])
def test_variant_mlem_vs_oracle(gpu_pkg, kernel, sens, path):
    pkg = gpu_pkg
    sc, ev = small(pkg)
    update = pkg.UPDATE_SENSITIVITY_DENOMINATOR if sens else pkg.UPDATE_CPU
    p = pkg.PATH_AUTO if path == "auto" else pkg.PATH_PER_EVENT
    with pkg.Reconstruction(sc, update=update, kernel=kernel, path=p) as rec:
        rec.set_events(ev)
        st = rec.stats()
        smap = rec.sensitivity(1)
        imgs = []
        for _ in range(3):
            rec(1)
            imgs.append(rec.rho)
    osens = ol.sensitivity_f32(sc, 1).astype(np.float64) if sens else None
    ref = ol.mlem_variant_f64(sc, ev, 3, kernel=kernel, sens=osens, n_threads=16)
    print(f"case: 64x64 @ 8 mm, kernel {kernel}, sensitivity {sens}, fused {st['fused']}",
          flush=True)
    assert st["fused"] == (1 if path == "auto" else 0)
    assert np.max(np.abs(smap - ol.sensitivity_f32(sc, 1))) < 2e-6
    for i in range(3):
        e = rel_l2(imgs[i], ref[i])
        print(f"iter {i + 1}: rel L2 vs FP64 {e:.2e}", flush=True)
        assert e < 1e-5
    if sens:
        # sum_p s_p rho'_p = number of events with a positive denominator
        lam = ol.denominators_f32(sc, ev, np.ones((64, 64), np.float32))
        assert (imgs[0].astype(np.float64) * smap).sum() == pytest.approx(
            np.count_nonzero(lam > 0), rel=1e-5)


def test_analytic_gradient_and_errors(gpu_pkg):
    """Image gradient of the log-likelihood with the analytic kernel: the
    MLEM identity (grad + 1) rho = rho'; the sigma derivative is refused."""
    pkg = gpu_pkg
    sc, ev = small(pkg, n=15000)
    rng = np.random.default_rng(9)
    rho = rng.uniform(0.5, 1.5, (64, 64)).astype(np.float32)
    with pkg.Reconstruction(sc, kernel=pkg.KERNEL_ANALYTIC) as rec:
        rec.set_events(ev)
        rec.rho = rho
        grad, _, _ = rec.backward(want_sigma=False)
        with pytest.raises(pkg.PetstripError):
            rec.backward(want_sigma=True)
        rec.rho = rho
        rec(1)
        nxt = rec.rho
    assert rel_l2((grad + 1) * rho, nxt) < 1e-5


@pytest.mark.parametrize("kernel", [1, 0])
def test_variant_fused_tall_boxes_vs_oracle(gpu_pkg, kernel):
    """The fused variant where the boxes take two row groups (128 x 128
    pixels of 2 mm: ~58 rows, one event per wave) and where they are packed
    (4 mm), two iterations against the FP64 oracle variant."""
    pkg = gpu_pkg
    for n_pix, size in ((128, 0.002), (96, 0.004)):
        sc = pkg.Scanner(0.43, 0.5, n_pix, n_pix, size, size, 0.01, 0.04)
        ev = pkg.phantom_events(sc, pkg.workloads.PHANTOM_3, 20000, seed=5)
        with pkg.Reconstruction(sc, update=pkg.UPDATE_SENSITIVITY_DENOMINATOR,
                                kernel=kernel) as rec:
            rec.set_events(ev)
            assert rec.stats()["fused"] == 1
            imgs = []
            for _ in range(2):
                rec(1)
                imgs.append(rec.rho)
        osens = ol.sensitivity_f32(sc, 1).astype(np.float64)
        ref = ol.mlem_variant_f64(sc, ev, 2, kernel=kernel, sens=osens, n_threads=16)
        for i in range(2):
            e = rel_l2(imgs[i], ref[i])
            print(f"{n_pix}x{n_pix} @ {size} m, kernel {kernel}, iteration {i + 1}: rel L2 "
                  f"vs FP64 {e:.2e}", flush=True)
            assert e < 1e-5


@pytest.mark.parametrize("force_bands", [1, 2])
def test_variant_gaussian_banded_vs_oracle(gpu_pkg, force_bands):
    """The Gaussian kernel with the sensitivity-weighted denominator on the
    banded walks (rho times the sensitivity in the band's pass 1, the
    band-combined denominators in pass 2), row bands and row x column
    bands, against the FP64 oracle variant."""
    pkg = gpu_pkg
    sc, ev = small(pkg)
    with pkg.Reconstruction(sc, update=pkg.UPDATE_SENSITIVITY_DENOMINATOR, kernel=0,
                            force_bands=force_bands) as rec:
        rec.set_events(ev)
        assert rec.stats()["fused"] == 1 + force_bands
        imgs = []
        for _ in range(2):
            rec(1)
            imgs.append(rec.rho)
    osens = ol.sensitivity_f32(sc, 1).astype(np.float64)
    ref = ol.mlem_variant_f64(sc, ev, 2, kernel=0, sens=osens, n_threads=16)
    for i in range(2):
        e = rel_l2(imgs[i], ref[i])
        print(f"bands {force_bands}, iteration {i + 1}: rel L2 vs FP64 {e:.2e}", flush=True)
        assert e < 1e-5
