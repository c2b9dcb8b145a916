"""Python host side of the 3D hybrid reconstruction (include/petstrip_hybrid.h).

HybridReconstruction mirrors the reference's PET3D::Hybrid::Reconstruction
as 3d_hybrid_reconstruction drives it (src/3d/hybrid/reconstruction_cmd.cpp:
456-640): construct from the geometry and the scanner's sigmas and length
(weights and sensitivity are formed then), add responses (frame events), call
it for n iterations, read rho (planes x rows x columns).
"""
import ctypes as C
import os

import numpy as np

from . import petstrip as _ps
from .hybrid_geometry import HybridGeometry, cylinder_responses, ring_geometry  # noqa: F401

GEOMETRY_WEIGHTS = 1
SENSITIVITY_ONE = 2

EVENT = np.dtype([("first", np.int16), ("second", np.int16), ("up", np.float32),
                  ("right", np.float32), ("tan", np.float32), ("sec", np.float32),
                  ("begin", np.int64), ("end", np.int64),
                  ("plane_begin", np.int16), ("plane_end", np.int16)], align=True)


class CHybridGeometry(C.Structure):
    _fields_ = [
        ("n_detectors", C.c_int32), ("n_columns", C.c_int32), ("n_rows", C.c_int32),
        ("n_planes", C.c_int32), ("pixel_size", C.c_double), ("lower_left_x", C.c_double),
        ("lower_left_y", C.c_double), ("z_left", C.c_double), ("n_pixel_infos", C.c_int64),
        ("segments", C.c_void_p), ("widths", C.c_void_p), ("info_begin", C.c_void_p),
        ("info_end", C.c_void_p), ("pixels", C.c_void_p), ("positions", C.c_void_p),
        ("weights", C.c_void_p),
    ]


class CHybridStats(C.Structure):
    _fields_ = [("n_events", C.c_int64), ("n_voxels", C.c_int64),
                ("last_iteration_ms", C.c_double), ("last_kernel_ms", C.c_double)]


OUTPUT_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_int32,
                        C.c_int32, C.c_int32)

_bound = False


def _lib():
    global _bound
    lib = _ps.load_library()
    if not _bound:
        vp, i64, i32, d = C.c_void_p, C.c_int64, C.c_int32, C.c_double
        sig = {
            "petstrip_hybrid_create": ([C.POINTER(CHybridGeometry), d, d, d, i32, i32,
                                        C.POINTER(vp)], C.c_int),
            "petstrip_hybrid_destroy": ([vp], None),
            "petstrip_hybrid_set_responses": ([vp, vp, vp, i64], C.c_int),
            "petstrip_hybrid_set_events": ([vp, vp, i64], C.c_int),
            "petstrip_hybrid_get_events": ([vp, vp], C.c_int),
            "petstrip_hybrid_set_weights": ([vp, vp, vp], C.c_int),
            "petstrip_hybrid_get_weights": ([vp, vp, vp], C.c_int),
            "petstrip_hybrid_set_sensitivity_3d": ([vp, vp, i32], C.c_int),
            "petstrip_hybrid_set_rho": ([vp, vp], C.c_int),
            "petstrip_hybrid_get_rho": ([vp, vp], C.c_int),
            "petstrip_hybrid_iterate": ([vp, i32], C.c_int),
            "petstrip_hybrid_run": ([vp, vp, i32, i32, i32, OUTPUT_CB, _ps.PROGRESS_CB, vp],
                                    C.c_int),
            "petstrip_hybrid_stats_get": ([vp, C.POINTER(CHybridStats)], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _bound = True
    return lib


def _check(rc):
    if rc != 0:
        raise _ps.PetstripError(rc, _ps.load_library().petstrip_last_error().decode())


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class HybridReconstruction:
    """The 3D hybrid reconstruction on one GPU (libpetstrip_hip.so).

    geometry: a HybridGeometry (GeometrySOA arrays and the voxel grid).
    flags: GEOMETRY_WEIGHTS keeps the geometry's weights, SENSITIVITY_ONE sets
    the sensitivity to one (the reference's --sens-to-one).
    """

    def __init__(self, geometry, sigma_z, sigma_dl, length, flags=0, device=0):
        lib = _lib()
        g = geometry
        self.geometry = g
        self._keep = [g.segments, g.widths, g.info_begin, g.info_end, g.pixels, g.positions,
                      g.weights]
        cg = CHybridGeometry(g.n_detectors, g.n_columns, g.n_rows, g.n_planes, g.pixel_size,
                             g.lower_left[0], g.lower_left[1], g.z_left, g.n_pixel_infos,
                             _p(g.segments), _p(g.widths), _p(g.info_begin), _p(g.info_end),
                             _p(g.pixels), _p(g.positions), _p(g.weights))
        self._h = C.c_void_p()
        _check(lib.petstrip_hybrid_create(C.byref(cg), sigma_z, sigma_dl, length, int(flags),
                                          int(device), C.byref(self._h)))

    def close(self):
        if self._h:
            _lib().petstrip_hybrid_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        g = self.geometry
        return (g.n_planes, g.n_rows, g.n_columns)

    def set_responses(self, lors, zzd):
        lors = np.ascontiguousarray(lors, np.int16).reshape(-1, 2)
        zzd = np.ascontiguousarray(zzd, np.float32).reshape(-1, 3)
        if len(lors) != len(zzd):
            raise ValueError("lors and zzd differ in length")
        _check(_lib().petstrip_hybrid_set_responses(self._h, _p(lors), _p(zzd), len(lors)))

    def set_events(self, events):
        events = np.ascontiguousarray(events, EVENT)
        _check(_lib().petstrip_hybrid_set_events(self._h, _p(events), len(events)))

    def events(self):
        out = np.zeros(self.stats()["n_events"], EVENT)
        _check(_lib().petstrip_hybrid_get_events(self._h, _p(out)))
        return out

    def weights(self):
        w = np.zeros(self.geometry.n_pixel_infos, np.float32)
        s = np.zeros(self.shape[1:], np.float32)
        _check(_lib().petstrip_hybrid_get_weights(self._h, _p(w), _p(s)))
        return w, s

    def set_weights(self, weights=None, sensitivity=None):
        w = None if weights is None else np.ascontiguousarray(weights, np.float32)
        s = None if sensitivity is None else np.ascontiguousarray(sensitivity, np.float32)
        _check(_lib().petstrip_hybrid_set_weights(self._h, _p(w), _p(s)))

    def set_sensitivity_3d(self, volume):
        """External 3D sensitivity (the command's --sensitivity,
        reconstruction_cmd.cpp:505-519): n_planes / 2 planes of rows x
        columns; the geometry weights are normalised by plane 0 and the
        denominator of plane iz takes plane |iz - n_planes / 2|."""
        v = np.ascontiguousarray(volume, np.float32)
        if v.ndim != 3 or v.shape[1:] != self.shape[1:]:
            raise ValueError(f"sensitivity volume of shape (D, rows, columns) expected, "
                             f"got {v.shape}")
        _check(_lib().petstrip_hybrid_set_sensitivity_3d(self._h, _p(v), v.shape[0]))

    def set_plane_weights(self, raw, sensitivity_mode=0, sensitivity=None):
        """Per-plane LOR-pixel weights, the reference's multi-plane geometry
        (one system matrix per half plane): raw (n_planes_half, n_pixel_infos);
        sensitivity_mode 0 sums each plane's weights per pixel, 1 sets one, 2
        takes `sensitivity` (n_planes_half, rows, columns); each plane's weights
        are divided by it, and voxel plane iz takes plane |iz - n_planes_half|
        (include/petstrip_hybrid.h)."""
        raw = np.ascontiguousarray(raw, np.float32)
        if raw.ndim == 1:
            raw = raw[None, :]
        sv = None if sensitivity is None else np.ascontiguousarray(sensitivity, np.float32)
        _check(_bind_system().petstrip_hybrid_set_plane_weights(
            self._h, _p(raw), raw.shape[0], int(sensitivity_mode), _p(sv)))

    def plane_weights(self):
        """The normalised weights (weight planes, n_pixel_infos) and the
        sensitivity (planes, rows, columns) the iteration uses: one plane
        each for the geometry weights and the 2D map."""
        lib = _bind_system()
        nw, ns = C.c_int32(), C.c_int32()
        _check(lib.petstrip_hybrid_get_plane_weights(self._h, None, None, C.byref(nw),
                                                     C.byref(ns)))
        w = np.zeros((nw.value, self.geometry.n_pixel_infos), np.float32)
        s = np.zeros((ns.value,) + self.shape[1:], np.float32)
        _check(lib.petstrip_hybrid_get_plane_weights(self._h, _p(w), _p(s), C.byref(nw),
                                                     C.byref(ns)))
        return w, s

    @property
    def rho(self):
        out = np.zeros(self.shape, np.float32)
        _check(_lib().petstrip_hybrid_get_rho(self._h, _p(out)))
        return out

    @rho.setter
    def rho(self, value):
        v = np.ascontiguousarray(value, np.float32).reshape(self.shape)
        _check(_lib().petstrip_hybrid_set_rho(self._h, _p(v)))

    def __call__(self, n=1):
        _check(_lib().petstrip_hybrid_iterate(self._h, int(n)))
        return self

    def run(self, n_blocks, n_iterations_in_block, rho=None, start_iteration=0,
            output=None, progress=None):
        """The reference's block schedule (cuda/reconstruction.cu:103-181):
        output(iteration, image) after iterations 1-5, 10, 15, 20, 30, 50, 100
        of the first block and after every block."""
        img = np.ones(self.shape, np.float32) if rho is None else \
            np.ascontiguousarray(rho, np.float32).reshape(self.shape).copy()

        def on_out(_u, it, ptr, w, h, d):
            if output is not None:
                output(it, np.ctypeslib.as_array(ptr, (d, h, w)).copy())
            return 0

        def on_prog(_u, done, fin):
            if progress is not None:
                progress(done, bool(fin))

        cb_o, cb_p = OUTPUT_CB(on_out), _ps.PROGRESS_CB(on_prog)
        _check(_lib().petstrip_hybrid_run(self._h, _p(img), n_blocks, n_iterations_in_block,
                                          start_iteration, cb_o, cb_p, None))
        return img

    def stats(self):
        s = CHybridStats()
        _check(_lib().petstrip_hybrid_stats_get(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in CHybridStats._fields_}


def _bind_system():
    lib = _lib()
    vp, i64, i32, d = C.c_void_p, C.c_int64, C.c_int32, C.c_double
    sig = {
        "petstrip_hybrid_set_plane_weights": ([vp, vp, i32, i32, vp], C.c_int),
        "petstrip_hybrid_get_plane_weights": ([vp, vp, vp, C.POINTER(C.c_int32),
                                               C.POINTER(C.c_int32)], C.c_int),
        "petstrip_system_matrix_read": ([C.c_char_p, C.POINTER(vp)], C.c_int),
        "petstrip_system_matrix_free": ([vp], None),
        "petstrip_system_matrix_info": ([vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
        "petstrip_system_matrix_weights": ([vp, C.POINTER(CHybridGeometry), vp], C.c_int),
        "petstrip_read_detector_centres": ([C.c_char_p, vp, i32, C.POINTER(C.c_int32)],
                                           C.c_int),
        "petstrip_hybrid_geometry_from_matrix": ([vp, vp, i32, i32, d, d, d, vp, i32, vp, vp,
                                                  vp, vp, vp, vp, vp, C.POINTER(C.c_int64)],
                                                 C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


def _cgeom(g):
    return CHybridGeometry(g.n_detectors, g.n_columns, g.n_rows, g.n_planes, g.pixel_size,
                           g.lower_left[0], g.lower_left[1], g.z_left, g.n_pixel_infos,
                           _p(g.segments), _p(g.widths), _p(g.info_begin), _p(g.info_end),
                           _p(g.pixels), _p(g.positions), _p(g.weights))


class SystemMatrix:
    """A 2D barrel system matrix file, full form (the reference's
    SparseMatrix, src/2d/barrel/sparse_matrix.h:163-238), read by the
    library (host only)."""

    def __init__(self, path):
        lib = _bind_system()
        self._m = C.c_void_p()
        _check(lib.petstrip_system_matrix_read(os.fsencode(path), C.byref(self._m)))
        npr, nd = C.c_int32(), C.c_int32()
        ne, nel = C.c_int64(), C.c_int64()
        _check(lib.petstrip_system_matrix_info(self._m, C.byref(npr), C.byref(nd),
                                               C.byref(ne), C.byref(nel)))
        self.n_pixels_in_row, self.n_detectors = npr.value, nd.value
        self.n_emissions, self.n_elements = ne.value, nel.value

    def close(self):
        if self._m:
            _bind_system().petstrip_system_matrix_free(self._m)
            self._m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def weights(self, g):
        """load_weights_from_matrix_file (geometry_soa.h:210-267) for geometry
        g: hits / emissions per pixel info (n_pixel_infos floats)."""
        w = np.zeros(g.n_pixel_infos, np.float32)
        cg = _cgeom(g)
        _check(_bind_system().petstrip_system_matrix_weights(self._m, C.byref(cg), _p(w)))
        return w

    def geometry(self, centres, pixel_size, n_planes=None, z_left=None, inactive=()):
        """The GeometrySOA of this matrix and the detector centres
        (geometry_soa.h:71-182) on the matrix's centred grid of pixel_size,
        the LORs of the inactive detectors empty; n_planes (default the
        matrix's pixels per row) and z_left (default centred) as the
        command line assumes them (reconstruction_cmd.cpp:380-397)."""
        lib = _bind_system()
        n = self.n_pixels_in_row
        n_planes = n if n_planes is None else int(n_planes)
        ps = np.float32(pixel_size)
        # PixelGrid(n, n, s): lower left (-s n / 2, -s n / 2) in float
        ll = float(np.float32(-ps * np.float32(n)) / np.float32(2))
        z_left = float(-n_planes * float(pixel_size) / 2) if z_left is None else float(z_left)
        cen = np.ascontiguousarray(centres, np.float32).reshape(-1, 2)
        ina = np.ascontiguousarray(inactive, np.int16)
        ni = C.c_int64()
        _check(lib.petstrip_hybrid_geometry_from_matrix(
            self._m, _p(cen), n, n, float(ps), ll, ll, _p(ina) if len(ina) else None, len(ina),
            None, None, None, None, None, None, None, C.byref(ni)))
        nd = self.n_detectors
        n_lors = nd * (nd + 1) // 2
        seg = np.zeros((n_lors, 4), np.float32)
        wid = np.zeros(n_lors, np.float32)
        beg = np.zeros(n_lors, np.int64)
        end = np.zeros(n_lors, np.int64)
        pix = np.zeros((max(ni.value, 1), 2), np.int16)
        pos = np.zeros(max(ni.value, 1), np.float32)
        w = np.zeros(max(ni.value, 1), np.float32)
        _check(lib.petstrip_hybrid_geometry_from_matrix(
            self._m, _p(cen), n, n, float(ps), ll, ll, _p(ina) if len(ina) else None, len(ina),
            _p(seg), _p(wid), _p(beg), _p(end), _p(pix), _p(pos), _p(w), C.byref(ni)))
        m = ni.value
        ring = float(np.mean(np.hypot(cen[:, 0], cen[:, 1]))) if len(cen) else 0.0
        return HybridGeometry(nd, n, n, float(ps), (ll, ll), z_left, n_planes, seg, wid, beg,
                              end, pix[:m], pos[:m], w[:m], ring)


def read_detector_centres(path):
    """Detector centres (n, 2) of a scanner description file
    (ScannerBuilder::deserialize, Polygon::center)."""
    lib = _bind_system()
    n = C.c_int32()
    _check(lib.petstrip_read_detector_centres(os.fsencode(path), None, 0, C.byref(n)))
    out = np.zeros((n.value, 2), np.float32)
    _check(lib.petstrip_read_detector_centres(os.fsencode(path), _p(out), n.value,
                                              C.byref(n)))
    return out
