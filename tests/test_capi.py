"""C ABI of libpetstrip_hip.so on the CPU: it loads, exports every function
include/petstrip.h declares, fails cleanly without a device, and its event
generator (host code) is seeded and thread-count independent."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "petstrip.h")


PROBE_HEADER = os.path.join(ROOT, "include", "petstrip_probe.h")


def declared_functions(header=HEADER):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(petstrip_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ["petstrip_create", "petstrip_destroy", "petstrip_set_events",
              "petstrip_iterate", "petstrip_run", "petstrip_forward",
              "petstrip_backward", "petstrip_last_error"]:
        assert n in names


HYBRID_HEADER = os.path.join(ROOT, "include", "petstrip_hybrid.h")


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.petstrip_abi_version() == 1
    hybrid = declared_functions(HYBRID_HEADER)
    assert "petstrip_hybrid_create" in hybrid and "petstrip_hybrid_run" in hybrid
    missing = [n for n in hybrid if not hasattr(lib, n)]
    assert not missing, missing


def test_hybrid_errors_are_returned_without_a_device(pkg):
    lib = pkg.load_library()
    lib.petstrip_hybrid_create.restype = C.c_int
    h = C.c_void_p()
    assert lib.petstrip_hybrid_create(None, C.c_double(0.01), C.c_double(0.04),
                                      C.c_double(0.3), 0, 0, C.byref(h)) == -1
    assert b"bad argument" in lib.petstrip_last_error()
    assert lib.petstrip_hybrid_iterate(None, 1) == -1


def test_probe_library_exports_every_declared_symbol(pkg):
    """libpetstrip_probe.so exports what include/petstrip_probe.h declares and
    rejects bad arguments before touching a device."""
    import ctypes
    lib = pkg.petstrip.load_probe_library()
    names = declared_functions(PROBE_HEADER)
    assert names == ["petstrip_probe_atomics"]
    assert all(hasattr(lib, n) for n in names)
    rate = ctypes.c_double(0.0)
    assert lib.petstrip_probe_atomics(0, 7, 512, 512, 16, ctypes.byref(rate), None) != 0
    assert lib.petstrip_probe_atomics(0, 0, 8, 512, 16, ctypes.byref(rate), None) != 0
    assert lib.petstrip_probe_atomics(0, 0, 512, 512, 0, ctypes.byref(rate), None) != 0


def test_errors_are_returned_not_aborted(pkg):
    lib = pkg.load_library()
    h = C.c_void_p()
    # invalid scanner
    sc = pkg.Scanner(0.43, 0.5, 1, 64, 0.01, 0.01, 0.01, 0.04).c()
    rc = lib.petstrip_create(C.byref(sc), 0, None, C.byref(h))
    assert rc == -1 and b"n_y_pixels" in lib.petstrip_last_error()
    sc = pkg.Scanner(0.43, 0.5, 64, 64, 0.01, 0.01, -1.0, 0.04).c()
    assert lib.petstrip_create(C.byref(sc), 0, None, C.byref(h)) == -1
    assert lib.petstrip_iterate(None, 1) == -1
    assert lib.petstrip_set_events(None, None, 0) == -1


def test_create_without_device_reports_edevice(pkg):
    if pkg.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(pkg.PetstripError) as e:
        pkg.Reconstruction(pkg.Scanner(0.43, 0.5, 64, 64, 0.01, 0.01, 0.01, 0.04))
    assert e.value.code == -2


def test_phantom_seeded_and_thread_independent(pkg):
    sc = pkg.Scanner(0.43, 0.5, 160, 160, 0.004, 0.004, 0.01, 0.04)
    a = pkg.phantom_events(sc, pkg.workloads.PHANTOM_3, 50000, seed=5, n_threads=1)
    b = pkg.phantom_events(sc, pkg.workloads.PHANTOM_3, 50000, seed=5, n_threads=7)
    c = pkg.phantom_events(sc, pkg.workloads.PHANTOM_3, 50000, seed=6, n_threads=7)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
    # prefix property: slot i does not depend on n
    d = pkg.phantom_events(sc, pkg.workloads.PHANTOM_3, 1000, seed=5)
    assert np.array_equal(a[:1000], d)


def test_phantom_geometry(pkg):
    """Noise-free geometry recovers the emission points inside the phantom;
    the hot region carries about intensity*area of the detected events and
    the cold region none."""
    sc = pkg.Scanner(0.43, 0.5, 160, 160, 0.004, 0.004, 1e-9, 1e-9)
    ev = pkg.phantom_events(sc, pkg.workloads.PHANTOM_3, 200000, seed=2).astype(np.float64)
    zu, zd, dl = ev.T
    t = (zu - zd) / (2 * 0.43)
    y = -0.5 * dl / np.sqrt(1 + t * t)
    z = 0.5 * (zu + zd + 2 * y * t)
    assert np.all(np.abs(zu) < 0.25 + 1e-6) and np.all(np.abs(zd) < 0.25 + 1e-6)
    inside_bg = (z / 0.20) ** 2 + (y / 0.15) ** 2 <= 1 + 1e-6
    assert inside_bg.mean() > 0.999
    cold = (z + 0.08) ** 2 + (y - 0.03) ** 2 < (0.03 - 1e-4) ** 2
    assert cold.sum() == 0
    hot = (z - 0.08) ** 2 + (y + 0.02) ** 2 < 0.04 ** 2
    # emission density ratio hot/background = 4 (detection efficiency varies
    # smoothly; allow 15 %)
    area_hot = np.pi * 0.04 ** 2
    area_bg = np.pi * 0.2 * 0.15 - area_hot - np.pi * 0.03 ** 2
    ratio = (hot.sum() / area_hot) / ((~hot & inside_bg).sum() / area_bg)
    assert 3.4 < ratio < 4.6


def test_phantom_rejects_bad_regions(pkg):
    sc = pkg.Scanner(0.43, 0.5, 16, 16, 0.01, 0.01, 0.01, 0.04)
    with pytest.raises(pkg.PetstripError):
        pkg.phantom_events(sc, [(0, 0, 0.1, 0.1, 0, 0.0)], 10)
    with pytest.raises(pkg.PetstripError):
        pkg.phantom_events(sc, [(0, 0, -0.1, 0.1, 0, 1.0)], 10)
