"""Event sets past 2^26 events: a launch of one wave per event over them
would need more than 2^32 work-items, which a dispatch cannot count, so those
launches go in slices.  Checked here at 7e7 events (generated on the device):
the per-event footprints of the last events against a handle holding only
them, and the per-event kernels (denominators, backprojection, sigma
derivative) against the binned fused kernel over the same events."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_LARGE = 70_000_000  # > 2^26 = 67,108,864
NAME = "strip-160px-4mm-1e6ev"


def _rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_footprints_past_2e26_events(gpu_pkg):
    pkg = gpu_pkg
    _, _, _, _, phantom, seed = pkg.WORKLOADS[NAME]
    sc = pkg.workload_scanner(NAME)
    tail = 3_000_000
    with pkg.Reconstruction(sc) as big:
        big.set_phantom_events(phantom, N_LARGE, seed=seed)
        nb, ne = big.footprints()
        st = big.stats()
        ev_tail = big.events()[-tail:].copy()
    print(f"case {NAME}: {N_LARGE} events, footprints of the last {tail}", flush=True)
    assert st["n_box_pixels"] == int(nb.astype(np.int64).sum())
    assert st["n_ellipse_pixels"] == int(ne.astype(np.int64).sum())
    with pkg.Reconstruction(sc) as small:
        small.set_events(ev_tail)
        nb_t, ne_t = small.footprints()
    assert np.array_equal(nb[-tail:], nb_t)
    assert np.array_equal(ne[-tail:], ne_t)
    assert ne_t.min() > 0


def test_per_event_kernels_past_2e26_events(gpu_pkg):
    pkg = gpu_pkg
    _, _, _, _, phantom, seed = pkg.WORKLOADS[NAME]
    sc = pkg.workload_scanner(NAME)
    tail = 3_000_000
    g = np.zeros(N_LARGE, np.float32)
    g[-tail:] = 1.0  # weights only on events past 2^26
    out = {}
    for path in (pkg.PATH_AUTO, pkg.PATH_PER_EVENT):
        # FP64 global sums on both sides (the per-event kernels always use
        # them), so the images differ by the kernels' own rounding only
        with pkg.Reconstruction(sc, path=path, accumulation=pkg.ACCUM_COMPENSATED) as rec:
            rec.set_phantom_events(phantom, N_LARGE, seed=seed)
            lam = rec.forward()
            adj, sig, _ = rec.backward(g)
            rec(1)
            out[path] = (lam, adj, np.asarray(sig), rec.rho.copy(), rec.stats()["fused"])
    print(f"case {NAME}: {N_LARGE} events, fused vs per-event", flush=True)
    f, p = out[pkg.PATH_AUTO], out[pkg.PATH_PER_EVENT]
    assert f[4] != 0 and p[4] == 0
    # denominators of every event, the tail included
    assert np.all(p[0][-tail:] > 0)
    assert np.max(np.abs(p[0] - f[0]) / f[0]) < 1e-4
    # adjoint and sigma derivative of the tail events alone
    assert _rel_l2(p[1], f[1]) < 1e-5
    assert np.all(np.abs(p[2] - f[2]) <= 1e-4 * np.abs(f[2]))
    # one MLEM iteration over all events
    assert _rel_l2(p[3], f[3]) < 1e-5
