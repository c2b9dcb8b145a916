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
        r.ref_strip_pixel_center.argtypes = [d, d, C.c_int, C.c_int, d, C.c_int,
                                             C.c_int, vp]
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


def ref_mlem(sc, events, n_iter, n_threads=1, rho0=None):
    r = ref()
    np_ = sc.n_y_pixels * sc.n_z_pixels
    ev = np.ascontiguousarray(events, dtype=np.float32).reshape(-1, 3)
    rho = np.ones(np_, np.float32) if rho0 is None else \
        np.ascontiguousarray(rho0, dtype=np.float32).reshape(-1).copy()
    iters = np.zeros((n_iter, np_), np.float32)
    assert sc.pixel_width == sc.pixel_height
    r.ref_strip_mlem(sc.radius, sc.scintillator_length, sc.n_y_pixels, sc.n_z_pixels,
                     sc.pixel_width, sc.sigma_z, sc.sigma_dl, P(ev), ev.shape[0],
                     P(rho), n_iter, n_threads, P(iters))
    return iters.reshape(n_iter, sc.n_y_pixels, sc.n_z_pixels)


# ---------------------------------------------------------------------------
# 3D hybrid reconstruction: oracle/hybrid_oracle.c and the reference harness
# oracle/_ref/libref_hybrid.so (PET3D::Hybrid::Reconstruction).
REF_HYBRID_SO = os.path.join(ROOT, "oracle", "_ref", "libref_hybrid.so")

HYBRID_EVENT = np.dtype([("first", np.int16), ("second", np.int16), ("up", np.float32),
                         ("right", np.float32), ("tan", np.float32), ("sec", np.float32),
                         ("begin", np.int64), ("end", np.int64),
                         ("plane_begin", np.int16), ("plane_end", np.int16)], align=True)


class OHybridGeometry(C.Structure):
    _fields_ = [
        ("n_detectors", C.c_int32), ("n_columns", C.c_int32), ("n_rows", C.c_int32),
        ("n_planes", C.c_int32), ("pixel_size", C.c_double), ("lower_left_x", C.c_double),
        ("lower_left_y", C.c_double), ("z_left", C.c_double), ("n_pixel_infos", C.c_int64),
        ("segments", C.c_void_p), ("widths", C.c_void_p), ("info_begin", C.c_void_p),
        ("info_end", C.c_void_p), ("pixels", C.c_void_p), ("positions", C.c_void_p),
    ]


def ohybrid(g):
    return OHybridGeometry(g.n_detectors, g.n_columns, g.n_rows, g.n_planes, g.pixel_size,
                           g.lower_left[0], g.lower_left[1], g.z_left, g.n_pixel_infos,
                           P(g.segments), P(g.widths), P(g.info_begin), P(g.info_end),
                           P(g.pixels), P(g.positions))


_h = None


def _hyb():
    global _h
    if _h is None:
        o = C.CDLL(ORACLE_SO)
        vp, i64, d = C.c_void_p, C.c_int64, C.c_double
        G = C.POINTER(OHybridGeometry)
        o.oracle_hybrid_weights_f32.argtypes = [G, C.c_int, vp, vp]
        o.oracle_hybrid_weights_f64.argtypes = [G, C.c_int, vp, vp]
        o.oracle_hybrid_frames_f32.argtypes = [G, d, d, vp, vp, vp, i64, vp]
        o.oracle_hybrid_frames_f32.restype = i64
        o.oracle_hybrid_mlem_f32.argtypes = [G, d, d, d, vp, vp, C.c_int, vp, i64, C.c_int,
                                             vp, vp]
        o.oracle_hybrid_mlem_f64.argtypes = [G, d, d, d, vp, vp, C.c_int, vp, i64, C.c_int,
                                             vp, vp]
        for t in ("f32", "f64"):
            getattr(o, f"oracle_hybrid_mlem_planes_{t}").argtypes = [
                G, d, d, d, vp, C.c_int, vp, C.c_int, vp, i64, C.c_int, vp, vp]
            getattr(o, f"oracle_hybrid_plane_weights_{t}").argtypes = [G, vp, C.c_int, C.c_int,
                                                                       vp, vp]
        _h = o
    return _h


def hybrid_weights(g, sensitivity_one=False, f64=False, external=None):
    """(weights, W x H sensitivity) after normalisation; external: a W x H
    map (an external volume's plane 0) that normalises the weights instead."""
    t = np.float64 if f64 else np.float32
    w = np.zeros(g.n_pixel_infos, t)
    if external is not None:
        s = np.array(external, t).reshape(g.n_rows, g.n_columns)
        mode = 2
    else:
        s = np.zeros((g.n_rows, g.n_columns), t)
        mode = int(sensitivity_one)
    og = ohybrid(g)
    fn = _hyb().oracle_hybrid_weights_f64 if f64 else _hyb().oracle_hybrid_weights_f32
    fn(C.byref(og), mode, P(w), P(s))
    return w, s


def hybrid_frames(g, sigma_z, sigma_dl, weights, lors, zzd):
    out = np.zeros(len(lors), HYBRID_EVENT)
    og = ohybrid(g)
    lors = np.ascontiguousarray(lors, np.int16)
    zzd = np.ascontiguousarray(zzd, np.float32)
    k = _hyb().oracle_hybrid_frames_f32(C.byref(og), sigma_z, sigma_dl, P(weights), P(lors),
                                        P(zzd), len(lors), P(out))
    return out[:k].copy()


def hybrid_mlem(g, sigma_z, sigma_dl, length, weights, sensitivity, events, n_iter,
                rho0=None, f64=False):
    """Images after every iteration, (n_iter, planes, rows, columns).  A 2D
    sensitivity (rows, columns) is the pixel's factor; a 3D one (D, rows,
    columns), D = planes / 2, is an external volume read at plane |iz - D|."""
    t = np.float64 if f64 else np.float32
    sensitivity = np.asarray(sensitivity)
    depth = sensitivity.shape[0] if sensitivity.ndim == 3 else 0
    shape = (g.n_planes, g.n_rows, g.n_columns)
    rho = np.ones(shape, t) if rho0 is None else np.ascontiguousarray(rho0, t)
    out = np.zeros((n_iter,) + shape, t)
    og = ohybrid(g)
    events = np.ascontiguousarray(events, HYBRID_EVENT)
    fn = _hyb().oracle_hybrid_mlem_f64 if f64 else _hyb().oracle_hybrid_mlem_f32
    fn(C.byref(og), sigma_z, sigma_dl, length, P(np.ascontiguousarray(weights, t)),
       P(np.ascontiguousarray(sensitivity, t)), depth, P(events), len(events), n_iter,
       P(rho), P(out))
    return out


def have_ref_hybrid():
    return os.path.exists(REF_HYBRID_SO)


def ref_hybrid(g, sigma_z, sigma_dl, length, lors, zzd, n_iter, sensitivity_one=False,
               n_threads=1, keep_images=True, sensitivity_3d=None):
    """The reference's own reconstruction on the same geometry and responses:
    dict of weights, sensitivity, frames (kept events: first, second, up,
    right, tan, sec, begin, end, plane begin, end), images per iteration and
    the iterations' wall time (n_threads OpenMP threads)."""
    r = C.CDLL(REF_HYBRID_SO)
    vp, i64, f = C.c_void_p, C.c_int64, C.c_float
    ci = C.c_int
    r.ref_hybrid_run.argtypes = ([ci, ci, ci, f, f, f, f, ci] + [vp] * 7 +
                                 [f, f, f, ci, vp, vp, i64, ci, vp, vp, vp, vp, vp, ci, vp,
                                  vp])
    r.ref_hybrid_run.restype = ci
    lors = np.ascontiguousarray(lors, np.int16)
    zzd = np.ascontiguousarray(zzd, np.float32)
    n = len(lors)
    w = np.zeros(g.n_pixel_infos, np.float32)
    s = np.zeros((g.n_rows, g.n_columns), np.float32)
    kept = C.c_int64(0)
    frames = np.zeros((n, 10), np.float64)
    rho = (np.zeros((n_iter, g.n_planes, g.n_rows, g.n_columns), np.float32)
           if keep_images else None)
    seconds = C.c_double(0.0)
    rc = r.ref_hybrid_run(g.n_detectors, g.n_columns, g.n_rows, g.pixel_size,
                          g.lower_left[0], g.lower_left[1], g.z_left, g.n_planes,
                          P(g.segments), P(g.widths), P(g.info_begin), P(g.info_end),
                          P(g.pixels), P(g.positions), P(g.weights), sigma_z, sigma_dl,
                          length, int(sensitivity_one), P(lors), P(zzd), n, n_iter, P(w),
                          P(s), C.byref(kept), P(frames), P(rho), int(n_threads),
                          C.byref(seconds),
                          None if sensitivity_3d is None else
                          P(np.ascontiguousarray(sensitivity_3d, np.float32)))
    if rc != 0:
        raise RuntimeError("reference hybrid harness failed")
    return {"weights": w, "sensitivity": s, "frames": frames[:kept.value], "rho": rho,
            "seconds": seconds.value}


def ref_hybrid_read_geometry(path):
    """The reference's reader of a binary geometry file, as GeometrySOA
    arrays (dict)."""
    r = C.CDLL(REF_HYBRID_SO)
    vp = C.c_void_p
    r.ref_hybrid_read_geometry.argtypes = [C.c_char_p, vp, vp, vp, vp, vp] + [vp] * 7
    r.ref_hybrid_read_geometry.restype = C.c_int
    nd, nc, nr = C.c_int(), C.c_int(), C.c_int()
    g3 = np.zeros(3, np.float32)
    ni = C.c_int64()
    args = [path.encode(), C.byref(nd), C.byref(nc), C.byref(nr), P(g3), C.byref(ni)]
    if r.ref_hybrid_read_geometry(*args, *([None] * 7)) != 0:
        raise RuntimeError("reference geometry reader failed")
    n_lors = nd.value * (nd.value + 1) // 2
    out = {"n_detectors": nd.value, "n_columns": nc.value, "n_rows": nr.value, "grid": g3,
           "segments": np.zeros((n_lors, 4), np.float32), "widths": np.zeros(n_lors, np.float32),
           "info_begin": np.zeros(n_lors, np.int64), "info_end": np.zeros(n_lors, np.int64),
           "pixels": np.zeros((ni.value, 2), np.int16),
           "positions": np.zeros(ni.value, np.float32), "weights": np.zeros(ni.value, np.float32)}
    keys = ["segments", "widths", "info_begin", "info_end", "pixels", "positions", "weights"]
    if r.ref_hybrid_read_geometry(*args, *[P(out[k]) for k in keys]) != 0:
        raise RuntimeError("reference geometry reader failed")
    return out


def hybrid_plane_weights(g, raw, sens_mode=0, sensitivity=None, f64=False):
    """(weights, sensitivity) of per-plane raw weights (planes, n_pixel_infos):
    the volume (planes, rows, columns) summed (0), one (1) or given (2), each
    plane's weights divided by it."""
    t = np.float64 if f64 else np.float32
    raw = np.ascontiguousarray(raw, t)
    planes = raw.shape[0]
    w = np.zeros_like(raw)
    sv = (np.zeros((planes, g.n_rows, g.n_columns), t) if sensitivity is None else
          np.array(sensitivity, t).reshape(planes, g.n_rows, g.n_columns))
    og = ohybrid(g)
    fn = getattr(_hyb(), f"oracle_hybrid_plane_weights_{'f64' if f64 else 'f32'}")
    fn(C.byref(og), P(raw), planes, int(sens_mode), P(w), P(sv))
    return w, sv


def hybrid_mlem_planes(g, sigma_z, sigma_dl, length, weights, sensitivity, events, n_iter,
                       rho0=None, f64=False):
    """Images after every iteration with per-plane weights (D, n_pixel_infos)
    and a sensitivity volume (D, rows, columns), voxel plane iz reading plane
    |iz - D| of both."""
    t = np.float64 if f64 else np.float32
    weights = np.ascontiguousarray(weights, t)
    sensitivity = np.ascontiguousarray(sensitivity, t)
    shape = (g.n_planes, g.n_rows, g.n_columns)
    rho = np.ones(shape, t) if rho0 is None else np.ascontiguousarray(rho0, t)
    out = np.zeros((n_iter,) + shape, t)
    og = ohybrid(g)
    events = np.ascontiguousarray(events, HYBRID_EVENT)
    fn = getattr(_hyb(), f"oracle_hybrid_mlem_planes_{'f64' if f64 else 'f32'}")
    fn(C.byref(og), sigma_z, sigma_dl, length, P(weights), weights.shape[0], P(sensitivity),
       sensitivity.shape[0], P(events), len(events), n_iter, P(rho), P(out))
    return out


def ref_hybrid_system(mode, g, sigma_z, sigma_dl, length, lors, zzd, n_iter,
                      matrix_paths, detector_file=None, inactive=(), sens_mode=0,
                      pixel_size=None, n_planes=None, z_left=None):
    """The reference's --system paths through oracle/_ref/libref_hybrid.so
    (ref_hybrid_run_system): mode 0 with geometry g and one matrix file per
    weight plane, mode 1 from one matrix file and a scanner description file
    (g ignored; pixel_size, n_planes, z_left give the grid).  Returns a dict
    with the geometry arrays (mode 1), weights (planes, n_pixel_infos),
    sensitivity, frames and the images after every iteration."""
    r = C.CDLL(REF_HYBRID_SO)
    vp, i64, f, ci = C.c_void_p, C.c_int64, C.c_float, C.c_int
    r.ref_hybrid_run_system.argtypes = (
        [ci, ci, ci, ci, f, f, f] + [vp] * 6 + [vp, ci, C.c_char_p, vp, ci, f, ci, f, f, f, ci,
                                                vp, vp, vp, i64, ci] + [vp] * 12)
    r.ref_hybrid_run_system.restype = ci
    paths = (C.c_char_p * len(matrix_paths))(*[os.fsencode(p) for p in matrix_paths])
    lors = np.ascontiguousarray(lors, np.int16)
    zzd = np.ascontiguousarray(zzd, np.float32)
    ina = np.ascontiguousarray(inactive, np.int16)
    if mode == 0:
        geo = [g.n_detectors, g.n_columns, g.n_rows, g.pixel_size, g.lower_left[0],
               g.lower_left[1]]
        arrays = [P(g.segments), P(g.widths), P(g.info_begin), P(g.info_end), P(g.pixels),
                  P(g.positions)]
        n_planes, z_left, nd, ncol = g.n_planes, g.z_left, g.n_detectors, g.n_columns
    else:
        geo = [0, 0, 0, pixel_size, 0.0, 0.0]
        arrays = [None] * 6
        nd = None
    n_infos, depth = C.c_int64(), C.c_int32()
    common = [float(z_left), int(n_planes), sigma_z, sigma_dl, length, int(sens_mode), None,
              P(lors), P(zzd), len(lors), n_iter]
    tail = [C.byref(n_infos), C.byref(depth)] + [None] * 10
    rc = r.ref_hybrid_run_system(mode, *geo, *arrays, paths, len(matrix_paths),
                                 None if detector_file is None else os.fsencode(detector_file),
                                 P(ina) if len(ina) else None, len(ina), *common, *tail)
    if rc != 0:
        raise RuntimeError("reference hybrid harness failed (sizes)")
    ni = n_infos.value
    if mode == 1:
        with open(matrix_paths[0], "rb") as fh:
            hdr = np.frombuffer(fh.read(16), "<u4")
        ncol, nd = int(hdr[1]), int(hdr[3])
    n_lors = nd * (nd + 1) // 2
    planes = len(matrix_paths) if mode == 0 else 1
    out = {"segments": np.zeros((n_lors, 4), np.float32),
           "begin": np.zeros(n_lors, np.int64), "end": np.zeros(n_lors, np.int64),
           "pixels": np.zeros((max(ni, 1), 2), np.int16),
           "positions": np.zeros(max(ni, 1), np.float32),
           "weights": np.zeros((planes, max(ni, 1)), np.float32)}
    np_ = ncol * ncol
    sens = np.zeros(np_ * max(planes, int(n_planes) // 2), np.float32)
    frames = np.zeros((len(lors), 10), np.float64)
    rho = np.zeros((n_iter, int(n_planes), ncol, ncol), np.float32)
    kept = C.c_int64()
    tail = [C.byref(n_infos), C.byref(depth), P(out["segments"]), P(out["begin"]),
            P(out["end"]), P(out["pixels"]), P(out["positions"]), P(out["weights"]), P(sens),
            C.byref(kept), P(frames), P(rho)]
    rc = r.ref_hybrid_run_system(mode, *geo, *arrays, paths, len(matrix_paths),
                                 None if detector_file is None else os.fsencode(detector_file),
                                 P(ina) if len(ina) else None, len(ina), *common, *tail)
    if rc != 0:
        raise RuntimeError("reference hybrid harness failed")
    out["n_pixel_infos"] = ni
    out["sensitivity"] = sens[:np_ * depth.value].reshape(depth.value, ncol, ncol)
    out["frames"] = frames[:kept.value]
    out["rho"] = rho
    return out
