"""petstrip_iterate_noalloc (include/petstrip.h): one MLEM iteration with
every buffer beyond its inputs and outputs in the library's static device
scratch array, through the C ABI, against the entry point whose kernels it
shares (petstrip_iterate on the per-event path) and against the oracle.
Two calls in a row, the second with a smaller input, so a stale scratch
accumulator would show; then an in-place call and the size limit.

Device memory for inputs and outputs comes from the HIP runtime the
library is linked with, called through ctypes (the test's plumbing); the
entry point itself allocates nothing."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def hip():
    rt = C.CDLL("libamdhip64.so.7")  # the runtime libpetstrip_hip.so loaded
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipSetDevice.argtypes = [C.c_int]
    return rt


def device_copy(rt, a):
    p = C.c_void_p()
    assert rt.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
    if a.nbytes:
        assert rt.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
    return p


def noalloc(pkg, sc, ev, rho, in_place=False):
    lib = pkg.load_library()
    lib.petstrip_iterate_noalloc.argtypes = [C.POINTER(pkg.petstrip.CScanner), C.c_void_p,
                                             C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.petstrip_iterate_noalloc.restype = C.c_int
    rt = hip()
    assert rt.hipSetDevice(0) == 0
    ev = np.ascontiguousarray(ev, np.float32)
    rho = np.ascontiguousarray(rho, np.float32)
    d_ev, d_rho = device_copy(rt, ev), device_copy(rt, rho)
    d_out = d_rho if in_place else device_copy(rt, np.zeros_like(rho))
    try:
        cs = sc.c()
        rc = lib.petstrip_iterate_noalloc(C.byref(cs), d_ev, len(ev), d_rho, d_out, None)
        if rc != 0:
            raise pkg.PetstripError(rc, lib.petstrip_last_error().decode())
        assert rt.hipDeviceSynchronize() == 0
        out = np.empty_like(rho)
        assert rt.hipMemcpy(out.ctypes.data, d_out, out.nbytes, D2H) == 0
        return out
    finally:
        for p in {d_ev.value, d_rho.value, d_out.value}:
            rt.hipFree(p)


def per_event(pkg, sc, ev, rho):
    with pkg.Reconstruction(sc, path=pkg.PATH_PER_EVENT) as rec:
        rec.set_events(ev)
        rec.rho = rho
        rec(1)
        return rec.rho


def test_noalloc_matches_iterate(gpu_pkg):
    pkg = gpu_pkg
    cases = [(64, 0.008, 30000, 3), (48, 0.009, 8000, 5)]  # the second smaller
    for n_pix, size, n, seed in cases:
        sc = pkg.Scanner(0.43, 0.5, n_pix, n_pix, size, size, 0.01, 0.04)
        ev = pkg.phantom_events(sc, pkg.workloads.PHANTOM_3, n, seed=seed)
        rho = np.random.default_rng(seed).uniform(0.5, 1.5, (n_pix, n_pix)).astype(np.float32)
        print(f"case: petstrip_iterate_noalloc {n_pix}x{n_pix}, {n} events", flush=True)
        got = noalloc(pkg, sc, ev, rho)
        want = per_event(pkg, sc, ev, rho)
        ref = ol.mlem_f32(sc, ev, 1, n_threads=8, rho0=rho)[0]
        e_api, e_ref = rel_l2(got, want), rel_l2(got, ref)
        print(f"  rel L2 vs petstrip_iterate (per-event path) {e_api:.2e}, "
              f"vs oracle {e_ref:.2e}", flush=True)
        # the same kernels: only the order of the FP64 atomic sums differs
        assert e_api < 1e-6
        assert e_ref < 1e-5
        assert np.all(got[-1, :] == 0) and np.all(got[:, -1] == 0)
    # in place (d_rho_out == d_rho)
    got2 = noalloc(pkg, sc, ev, rho, in_place=True)
    assert rel_l2(got2, want) < 1e-6


def test_noalloc_size_limit(gpu_pkg):
    pkg = gpu_pkg
    lib = pkg.load_library()
    lib.petstrip_noalloc_max_events.argtypes = [C.c_int32, C.c_int32, C.POINTER(C.c_int64)]
    m = C.c_int64(0)
    assert lib.petstrip_noalloc_max_events(64, 64, C.byref(m)) == 0
    assert 1_000_000 < m.value < (64 << 20) // 56
    sc = pkg.Scanner(0.43, 0.5, 64, 64, 0.008, 0.008, 0.01, 0.04)
    ev = np.zeros((m.value + 1, 3), np.float32)
    with pytest.raises(pkg.PetstripError, match="exceed"):
        noalloc(pkg, sc, ev, np.ones((64, 64), np.float32))
