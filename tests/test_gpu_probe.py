"""The atomic-throughput probes and the binned kernel's atomic counts.

Every ellipse pixel's update in pass 2 is an LDS atomic; the flush makes at
most one global atomic per window pixel (petstrip_stats.n_window_pixels),
where the reference's GPU path makes one global atomic per ellipse pixel
(src/2d/strip/cuda/warp_granularity.cuh:145-150)."""
import math

import pytest

pytestmark = pytest.mark.gpu


def test_probe_rates(gpu_pkg):
    pkg = gpu_pkg
    for kind in (pkg.PROBE_GLOBAL_F32, pkg.PROBE_GLOBAL_F64, pkg.PROBE_GLOBAL_U64,
                 pkg.PROBE_LDS_U64, pkg.PROBE_LDS_U32, pkg.PROBE_LDS_F32):
        print(f"case: probe kind {kind}", flush=True)
        rate, ms = pkg.probe_atomics(kind, 256, 256, reps=64)
        print(f"  {rate:.3e} atomics/s in {ms:.3f} ms", flush=True)
        assert math.isfinite(rate) and rate > 1e9 and ms > 0.0


def test_window_pixels_bound_global_atomics(gpu_pkg):
    pkg = gpu_pkg
    name = "strip-160px-4mm-1e6ev"
    _, _, _, _, phantom, seed = pkg.WORKLOADS[name]
    sc = pkg.workload_scanner(name)
    with pkg.Reconstruction(sc) as rec:
        rec.set_phantom_events(phantom, 200000, seed=seed)
        st = rec.stats()
    print(f"case: {name} 2e5 events, fused {st['fused']}, window pixels "
          f"{st['n_window_pixels']}, ellipse pixels {st['n_ellipse_pixels']}", flush=True)
    assert st["fused"] == 1
    # windows of at most 1024 events: far fewer flushes than pixel updates
    assert 0 < st["n_window_pixels"] < st["n_ellipse_pixels"] / 4
    with pkg.Reconstruction(sc, path=pkg.PATH_PER_EVENT) as rec:
        rec.set_phantom_events(phantom, 1000, seed=seed)
        assert rec.stats()["n_window_pixels"] == 0
