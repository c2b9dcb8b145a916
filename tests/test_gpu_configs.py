"""The two large workloads, on the GPU, against the CPU oracle and through
size-independent properties.

- 512 x 512 pixels of 1 mm, 2e7 events of the six-ellipse phantom, 10 MLEM
  iterations (the headline workload): 10 iterations over a 1e6-event prefix
  against the FP32 oracle (bitwise equal to the reference CPU build, 1e-5
  relative L2 per iteration) and the FP64 oracle (1e-4); then all 2e7
  events for 10 iterations with the invariants of every iteration:
  sum rho' = the number of events with a positive denominator, the last row
  and column zero, and the int64 fixed-point accumulation bit-identical run
  to run.
- 1024 x 1024 pixels of 0.5 mm, 1e8 events of the ten-ellipse phantom (the
  calibration workload; its boxes are ~234 rows tall, so every launch runs
  the row-banded passes): two iterations and the log-likelihood gradient
  over a 2e5-event prefix against both oracles, then one iteration and the
  gradient over all 1e8 events with the sum invariant and the MLEM identity
  (grad + 1) rho = rho' of the next iteration.

The oracle threads follow the CPU share the process may use.
"""
import os

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def gpu_iterations(pkg, sc, ev, n_iter, **kw):
    with pkg.Reconstruction(sc, **kw) as rec:
        rec.set_events(ev)
        out = []
        for _ in range(n_iter):
            rec(1)
            out.append(rec.rho)
    return out


W512 = "strip-512px-1mm-2e7ev"
W1024 = "strip-1024px-0.5mm-1e8ev"


@pytest.mark.parametrize("oracle", ["f32", "f64"])
def test_512px_prefix_10_iterations(gpu_pkg, oracle):
    pkg = gpu_pkg
    sc = pkg.workload_scanner(W512)
    ev = pkg.workload_events(W512, n=1_000_000)
    print(f"case: {W512} first 1e6 events, 10 iterations vs {oracle} oracle "
          f"({THREADS} threads)", flush=True)
    got = gpu_iterations(pkg, sc, ev, 10)
    if oracle == "f32":
        ref, tol = ol.mlem_f32(sc, ev, 10, n_threads=THREADS), 1e-5
    else:
        ref, tol = ol.mlem_f64(sc, ev, 10, n_threads=THREADS), 1e-4
    for i in range(10):
        e = rel_l2(got[i], ref[i])
        print(f"  iteration {i + 1}: rel L2 {e:.2e}", flush=True)
        assert e < tol


def test_512px_width_classes(gpu_pkg):
    """At 512 x 512 pixels of 1 mm the widest windows (the boxes of steep
    events, ~95 columns) need more LDS than two workgroups per CU leave, so
    the narrower boxes bin apart: some work items, not all, fit two per CU,
    and the iteration over both kinds of launch matches the FP32 oracle."""
    pkg = gpu_pkg
    sc = pkg.workload_scanner(W512)
    ev = pkg.workload_events(W512, n=200_000)
    with pkg.Reconstruction(sc) as rec:
        rec.set_events(ev)
        st = rec.stats()
        print(f"case: {W512} first 2e5 events: {st['n_items_two_per_cu']} of "
              f"{st['n_work_items']} items two per CU", flush=True)
        assert 0 < st["n_items_two_per_cu"] < st["n_work_items"]
        got = []
        for _ in range(2):
            rec(1)
            got.append(rec.rho)
    ref = ol.mlem_f32(sc, ev, 2, n_threads=THREADS)
    for i in range(2):
        e = rel_l2(got[i], ref[i])
        print(f"  iteration {i + 1}: rel L2 {e:.2e}", flush=True)
        assert e < 1e-5


def _sum_invariant(img, rec, n):
    st = rec.stats()
    n_pos = n - st["n_zero_denominator"]
    s = float(np.sum(img, dtype=np.float64))
    return s, n_pos, abs(s - n_pos) / n_pos


def test_512px_full_10_iterations_properties(gpu_pkg):
    pkg = gpu_pkg
    n_pix, size, n_ev, _, phantom, seed = pkg.WORKLOADS[W512]
    sc = pkg.workload_scanner(W512)
    runs = []
    for acc in (pkg.ACCUM_FIXED64, pkg.ACCUM_FIXED64, pkg.ACCUM_FP32):
        print(f"case: {W512} all 2e7 events, 10 iterations, accumulation {acc}",
              flush=True)
        with pkg.Reconstruction(sc, accumulation=acc) as rec:
            rec.set_phantom_events(phantom, n_ev, seed=seed)
            imgs = []
            for i in range(10):
                rec(1)
                img = rec.rho
                s, n_pos, err = _sum_invariant(img, rec, n_ev)
                print(f"  iteration {i + 1}: sum rho' {s:.6f}, events with lambda > 0 "
                      f"{n_pos}, rel {err:.1e}", flush=True)
                assert err < 1e-5
                assert np.all(img[-1, :] == 0) and np.all(img[:, -1] == 0)
                assert np.all(img >= 0)
                imgs.append(img)
        runs.append(imgs)
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    e = rel_l2(runs[2][-1], runs[0][-1])
    print(f"  FP32 against FIXED64 after 10 iterations: {e:.2e}", flush=True)
    assert e < 1e-5


def test_1024px_prefix_iterations_and_gradient(gpu_pkg):
    pkg = gpu_pkg
    sc = pkg.workload_scanner(W1024)
    ev = pkg.workload_events(W1024, n=200_000)
    print(f"case: {W1024} first 2e5 events, 2 iterations + gradient ({THREADS} threads)",
          flush=True)
    with pkg.Reconstruction(sc) as rec:
        rec.set_events(ev)
        assert rec.stats()["fused"] in (2, 3)  # the row-banded passes
        got = []
        for _ in range(2):
            rec(1)
            got.append(rec.rho)
        lam = rec.forward()
        grad, gsig, ll = rec.backward(want_loglik=True)
    r32 = ol.mlem_f32(sc, ev, 2, n_threads=THREADS)
    r64 = ol.mlem_f64(sc, ev, 2, n_threads=THREADS)
    for i in range(2):
        e32, e64 = rel_l2(got[i], r32[i]), rel_l2(got[i], r64[i])
        print(f"  iteration {i + 1}: rel L2 vs f32 {e32:.2e}, vs f64 {e64:.2e}", flush=True)
        assert e32 < 1e-5 and e64 < 1e-4
    ll_ref, lam_ref, g_ref, gs_ref = ol.loglik_f64(sc, ev, got[1], n_threads=THREADS,
                                                   want_lambda=True)
    print(f"  lambda {rel_l2(lam, lam_ref):.2e} grad {rel_l2(grad, g_ref):.2e} sigma "
          f"{gsig} vs {tuple(gs_ref)} loglik {ll} vs {ll_ref}", flush=True)
    assert rel_l2(lam, lam_ref) < 2e-6
    assert rel_l2(grad, g_ref) < 1e-5
    assert gsig[0] == pytest.approx(gs_ref[0], rel=2e-4)
    assert gsig[1] == pytest.approx(gs_ref[1], rel=2e-4)
    assert ll == pytest.approx(ll_ref, rel=1e-6)


def test_1024px_full_iteration_and_gradient(gpu_pkg):
    pkg = gpu_pkg
    n_pix, size, n_ev, _, phantom, seed = pkg.WORKLOADS[W1024]
    sc = pkg.workload_scanner(W1024)
    print(f"case: {W1024} all 1e8 events, 1 iteration, gradient, 1 iteration", flush=True)
    with pkg.Reconstruction(sc) as rec:
        rec.set_phantom_events(phantom, n_ev, seed=seed)
        rec(1)
        rho1 = rec.rho
        s, n_pos, err = _sum_invariant(rho1, rec, n_ev)
        print(f"  iteration 1: sum rho' {s:.3f}, events with lambda > 0 {n_pos}, "
              f"rel {err:.1e}", flush=True)
        assert err < 1e-5
        assert np.all(rho1[-1, :] == 0) and np.all(rho1[:, -1] == 0)
        grad, _, _ = rec.backward()
        rec(1)
        rho2 = rec.rho
    # MLEM identity: the log-likelihood gradient at rho is sum_e K / lambda_e
    # - 1, and one MLEM iteration gives rho' = rho * sum_e K / lambda_e
    e = rel_l2((grad.astype(np.float64) + 1.0) * rho1, rho2)
    print(f"  (grad + 1) rho against the next iteration: {e:.2e}", flush=True)
    assert e < 1e-5
