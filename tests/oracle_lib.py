"""ctypes access to the CPU oracle (oracle/liboracle_strip.so) and, where it
was built, the reference harness (oracle/_ref/libref_strip.so).

Test infrastructure: only tests, the smoke check and the benchmark's CPU
baseline use it.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle_strip.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref_strip.so")


class OScanner(C.Structure):
    _fields_ = [
        ("radius", C.c_double), ("scintillator_length", C.c_double),
        ("n_y_pixels", C.c_int32), ("n_z_pixels", C.c_int32),
        ("pixel_height", C.c_double), ("pixel_width", C.c_double),
        ("sigma_z", C.c_double), ("sigma_dl", C.c_double),
        ("center_y", C.c_double), ("center_z", C.c_double),
    ]


_o = None
_r = None


def P(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def oracle():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            raise ImportError(f"{ORACLE_SO} missing: run make -C oracle")
        o = C.CDLL(ORACLE_SO)
        vp, i64 = C.c_void_p, C.c_int64
        S = C.POINTER(OScanner)
        o.oracle_mlem_f32.argtypes = [S, vp, i64, vp, C.c_int, C.c_int, vp, vp]
        o.oracle_mlem_f64.argtypes = [S, vp, i64, vp, C.c_int, C.c_int, vp]
        o.oracle_footprint.argtypes = [S, vp, i64, vp, vp, C.c_int]
        o.oracle_denominators_f32.argtypes = [S, vp, i64, vp, vp, C.c_int]
        o.oracle_kernel_f32.argtypes = [C.c_double, C.c_double] + [C.c_float] * 6
        o.oracle_kernel_f32.restype = C.c_float
        o.oracle_kernel_f64.argtypes = [C.c_double] * 8
        o.oracle_kernel_f64.restype = C.c_double
        o.oracle_ellipse_bb_f32.argtypes = [C.c_double, C.c_double, C.c_float, vp]
        o.oracle_ellipse_bb_f64.argtypes = [C.c_double, C.c_double, C.c_double, vp]
        o.oracle_tan_y_z_f32.argtypes = [C.c_double, vp, vp]
        o.oracle_tan_y_z_f64.argtypes = [C.c_double, vp, vp]
        o.oracle_pixel_at_f32.argtypes = [S, C.c_float, C.c_float, vp]
        o.oracle_pixel_center_f32.argtypes = [S, C.c_int32, C.c_int32, vp]
        o.oracle_sensitivity_f32.argtypes = [S, C.c_int, vp]
        o.oracle_sensitivity_point_f64.argtypes = [C.c_double] * 4
        o.oracle_sensitivity_point_f64.restype = C.c_double
        o.oracle_loglik_f64.argtypes = [S, vp, i64, vp, vp, C.c_double, C.c_double,
                                        C.c_double, C.c_double, vp, vp, vp, vp, C.c_int]
        o.oracle_analytic_kernel_f32.argtypes = [C.c_double, C.c_double] + [C.c_float] * 6
        o.oracle_analytic_kernel_f32.restype = C.c_float
        o.oracle_analytic_kernel_f64.argtypes = [C.c_double] * 8
        o.oracle_analytic_kernel_f64.restype = C.c_double
        o.oracle_mlem_variant_f64.argtypes = [S, vp, i64, vp, C.c_int, C.c_int, vp,
                                              C.c_int, vp]
        _o = o
    return _o


def have_ref():
    return os.path.exists(REF_SO)


def ref():
    global _r
    if _r is None:
        r = C.CDLL(REF_SO)
        vp, i64, d = C.c_void_p, C.c_int64, C.c_double
        r.ref_strip_mlem.argtypes = [d, d, C.c_int, C.c_int, d, d, d, vp, i64, vp,
                                     C.c_int, C.c_int, vp]
        r.ref_strip_sensitivity.argtypes = [d, d, C.c_int, C.c_int, d, d, d, vp]
        r.ref_strip_pixel_sensitivity.argtypes = [d, d, C.c_int, C.c_int, d, d, d, vp]
        r.ref_strip_kernel_f32.argtypes = [d] * 8
        r.ref_strip_kernel_f32.restype = d
        r.ref_strip_kernel_f64.argtypes = [d] * 8
        r.ref_strip_kernel_f64.restype = d
        r.ref_strip_ellipse_bb.argtypes = [d, d, d, vp]
        r.ref_strip_analytic_kernel_f32.argtypes = [d] * 8
        r.ref_strip_analytic_kernel_f32.restype = d
        r.ref_strip_analytic_kernel_f64.argtypes = [d] * 8
        r.ref_strip_analytic_kernel_f64.restype = d
        r.ref_strip_tan_y_z.argtypes = [d, vp, vp]
        r.ref_strip_pixel_at.argtypes = [d, d, C.c_int, C.c_int, d, C.c_float,
                                         C.c_float, vp]
        _r = r
    return _r


def oscanner(sc):
    """Accepts anything with the Scanner attributes."""
    return OScanner(sc.radius, sc.scintillator_length, sc.n_y_pixels, sc.n_z_pixels,
                    sc.pixel_height, sc.pixel_width, sc.sigma_z, sc.sigma_dl,
                    getattr(sc, "center_y", 0.0), getattr(sc, "center_z", 0.0))


class SimpleScanner:
    def __init__(self, radius, length, n_y, n_z, ph, pw, sz, sdl, cy=0.0, cz=0.0):
        self.radius, self.scintillator_length = radius, length
        self.n_y_pixels, self.n_z_pixels = n_y, n_z
        self.pixel_height, self.pixel_width = ph, pw
        self.sigma_z, self.sigma_dl = sz, sdl
        self.center_y, self.center_z = cy, cz


def mlem_f32(sc, events, n_iter, n_threads=1, rho0=None, counts=False):
    o = oracle()
    s = oscanner(sc)
    np_ = sc.n_y_pixels * sc.n_z_pixels
    ev = np.ascontiguousarray(events, dtype=np.float32).reshape(-1, 3)
    rho = np.ones(np_, np.float32) if rho0 is None else \
        np.ascontiguousarray(rho0, dtype=np.float32).reshape(-1).copy()
    iters = np.zeros((n_iter, np_), np.float32)
    cnt = np.zeros(2, np.uint64)
    o.oracle_mlem_f32(C.byref(s), P(ev), ev.shape[0], P(rho), n_iter, n_threads,
                      P(iters), P(cnt))
    iters = iters.reshape(n_iter, sc.n_y_pixels, sc.n_z_pixels)
    return (iters, cnt) if counts else iters


def mlem_f64(sc, events, n_iter, n_threads=1, rho0=None):
    o = oracle()
    s = oscanner(sc)
    np_ = sc.n_y_pixels * sc.n_z_pixels
    ev = np.ascontiguousarray(events, dtype=np.float32).reshape(-1, 3)
    rho = np.ones(np_, np.float64) if rho0 is None else \
        np.ascontiguousarray(rho0, dtype=np.float64).reshape(-1).copy()
    iters = np.zeros((n_iter, np_), np.float64)
    o.oracle_mlem_f64(C.byref(s), P(ev), ev.shape[0], P(rho), n_iter, n_threads,
                      P(iters))
    return iters.reshape(n_iter, sc.n_y_pixels, sc.n_z_pixels)


def mlem_variant_f64(sc, events, n_iter, kernel=1, sens=None, n_threads=8, rho0=None):
    """The reference --gpu iteration in FP64: kernel 0 = Gaussian,
    1 = analytic; sens = the map in the denominator (None: 1)."""
    o = oracle()
    s = oscanner(sc)
    np_ = sc.n_y_pixels * sc.n_z_pixels
    ev = np.ascontiguousarray(events, dtype=np.float32).reshape(-1, 3)
    rho = np.ones(np_, np.float64) if rho0 is None else \
        np.ascontiguousarray(rho0, dtype=np.float64).reshape(-1).copy()
    sv = None if sens is None else np.ascontiguousarray(sens, np.float64).reshape(-1)
    iters = np.zeros((n_iter, np_), np.float64)
    o.oracle_mlem_variant_f64(C.byref(s), P(ev), ev.shape[0], P(rho), n_iter, kernel,
                              P(sv), n_threads, P(iters))
    return iters.reshape(n_iter, sc.n_y_pixels, sc.n_z_pixels)


def footprint(sc, events, n_threads=8):
    o = oracle()
    s = oscanner(sc)
    ev = np.ascontiguousarray(events, dtype=np.float32).reshape(-1, 3)
    nb = np.zeros(ev.shape[0], np.int32)
    ne = np.zeros(ev.shape[0], np.int32)
    o.oracle_footprint(C.byref(s), P(ev), ev.shape[0], P(nb), P(ne), n_threads)
    return nb, ne


def denominators_f32(sc, events, rho, n_threads=8):
    o = oracle()
    s = oscanner(sc)
    ev = np.ascontiguousarray(events, dtype=np.float32).reshape(-1, 3)
    r = np.ascontiguousarray(rho, dtype=np.float32).reshape(-1)
    d = np.zeros(ev.shape[0], np.float32)
    o.oracle_denominators_f32(C.byref(s), P(ev), ev.shape[0], P(r), P(d), n_threads)
    return d


def loglik_f64(sc, events, rho, sk=None, ss=None, sens=None, n_threads=8,
               want_lambda=False, want_grad=True):
    o = oracle()
    s = oscanner(sc)
    sk = sk or (sc.sigma_z, sc.sigma_dl)
    ss = ss or sk
    ev = np.ascontiguousarray(events, dtype=np.float32).reshape(-1, 3)
    r = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
    ll = np.zeros(1, np.float64)
    lam = np.zeros(ev.shape[0], np.float64) if want_lambda else None
    g = np.zeros(r.shape[0], np.float64) if want_grad else None
    gs = np.zeros(2, np.float64)
    o.oracle_loglik_f64(C.byref(s), P(ev), ev.shape[0], P(r),
                        P(None if sens is None else np.ascontiguousarray(sens, np.float64)),
                        sk[0], sk[1], ss[0], ss[1], P(ll), P(lam), P(g), P(gs), n_threads)
    return ll[0], lam, (None if g is None else g.reshape(sc.n_y_pixels, sc.n_z_pixels)), gs


def sensitivity_f32(sc, mode):
    o = oracle()
    s = oscanner(sc)
    out = np.zeros((sc.n_y_pixels, sc.n_z_pixels), np.float32)
    o.oracle_sensitivity_f32(C.byref(s), mode, P(out))
    return out


def ref_mlem(sc, events, n_iter, n_threads=1):
    r = ref()
    np_ = sc.n_y_pixels * sc.n_z_pixels
    ev = np.ascontiguousarray(events, dtype=np.float32).reshape(-1, 3)
    rho = np.ones(np_, np.float32)
    iters = np.zeros((n_iter, np_), np.float32)
    assert sc.pixel_width == sc.pixel_height
    r.ref_strip_mlem(sc.radius, sc.scintillator_length, sc.n_y_pixels, sc.n_z_pixels,
                     sc.pixel_width, sc.sigma_z, sc.sigma_dl, P(ev), ev.shape[0],
                     P(rho), n_iter, n_threads, P(iters))
    return iters.reshape(n_iter, sc.n_y_pixels, sc.n_z_pixels)
