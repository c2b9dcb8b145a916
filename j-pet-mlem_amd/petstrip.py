"""Python host side of petstrip: ctypes binding of libpetstrip_hip.so.

Mirrors the reference's strip reconstruction interface
(grey277/j-pet-mlem):
  Scanner(radius, scintillator_length, n_y_pixels, n_z_pixels, pixel_height,
          pixel_width, sigma_z, sigma_dl, center_y=0, center_z=0)
                                            src/2d/strip/scanner.h:39-49
  Reconstruction(scanner): rho = 1, responses, operator()(n_iterations)
                                            src/2d/strip/reconstruction.h:57-154
  GPU::Reconstruction::run(...)             src/2d/strip/cuda/reconstruction.h:31-44
Every compute call goes through the HIP library; there is no CPU fallback.
A failing call raises PetstripError with the library's message.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpetstrip_hip.so")

ACCUM_FP32, ACCUM_COMPENSATED, ACCUM_FIXED64 = 0, 1, 2
UPDATE_CPU, UPDATE_SENSITIVITY_DENOMINATOR = 0, 1
PATH_AUTO, PATH_PER_EVENT = 0, 1
KERNEL_GAUSSIAN, KERNEL_ANALYTIC = 0, 1


class PetstripError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"petstrip error {code}: {msg}")
        self.code = code


class CScanner(C.Structure):
    _fields_ = [
        ("radius", C.c_double), ("scintillator_length", C.c_double),
        ("n_y_pixels", C.c_int32), ("n_z_pixels", C.c_int32),
        ("pixel_height", C.c_double), ("pixel_width", C.c_double),
        ("sigma_z", C.c_double), ("sigma_dl", C.c_double),
        ("center_y", C.c_double), ("center_z", C.c_double),
    ]


class COptions(C.Structure):
    _fields_ = [("accumulation", C.c_int32), ("update", C.c_int32),
                ("path", C.c_int32), ("row_scan", C.c_int32),
                ("force_bands", C.c_int32), ("kernel", C.c_int32),
                ("row_table", C.c_int32), ("pass1_walk", C.c_int32)]


class CPhantomGrid(C.Structure):
    _fields_ = [("n_columns", C.c_int32), ("n_rows", C.c_int32),
                ("pixel_size", C.c_double), ("lower_left_x", C.c_double),
                ("lower_left_y", C.c_double), ("tan_bins", C.c_int32),
                ("max_tan", C.c_double)]


class CPhantomOutput(C.Structure):
    _fields_ = [("responses", C.c_void_p), ("exact", C.c_void_p), ("events", C.c_void_p),
                ("emitted", C.c_void_p), ("detected", C.c_void_p), ("naive", C.c_void_p),
                ("tan_bins_map", C.c_void_p), ("n_responses", C.c_int64),
                ("n_emitted", C.c_int64)]


class CStats(C.Structure):
    _fields_ = [
        ("n_events", C.c_int64), ("n_box_pixels", C.c_int64),
        ("n_ellipse_pixels", C.c_int64), ("n_zero_denominator", C.c_int64),
        ("flop_alg", C.c_double), ("last_iteration_ms", C.c_double),
        ("last_kernel_ms", C.c_double), ("fused", C.c_int32),
        ("n_work_items", C.c_int32), ("lds_bytes", C.c_int64),
        ("n_window_pixels", C.c_int64), ("row_table_bytes", C.c_int64),
        ("horner_walks", C.c_int64), ("n_items_two_per_cu", C.c_int64),
    ]


OUTPUT_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_float),
                        C.c_int32, C.c_int32)
PROGRESS_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int)

_lib = None


def load_library(path=LIB_PATH):
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(path)
    vp, i32, i64, dp, fp = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_void_p
    sig = {
        "petstrip_abi_version": ([], C.c_int),
        "petstrip_last_error": ([], C.c_char_p),
        "petstrip_device_count": ([C.POINTER(C.c_int)], C.c_int),
        "petstrip_create": ([C.POINTER(CScanner), C.c_int, C.POINTER(COptions),
                             C.POINTER(vp)], C.c_int),
        "petstrip_destroy": ([vp], C.c_int),
        "petstrip_device_name": ([vp, C.c_char_p, C.c_size_t], C.c_int),
        "petstrip_stream": ([vp, C.POINTER(vp)], C.c_int),
        "petstrip_set_options": ([vp, C.POINTER(COptions)], C.c_int),
        "petstrip_set_events": ([vp, fp, i64], C.c_int),
        "petstrip_set_events_device": ([vp, vp, i64], C.c_int),
        "petstrip_set_rho": ([vp, fp], C.c_int),
        "petstrip_get_rho": ([vp, fp], C.c_int),
        "petstrip_set_rho_device": ([vp, vp], C.c_int),
        "petstrip_rho_device": ([vp, C.POINTER(vp)], C.c_int),
        "petstrip_iterate": ([vp, C.c_int], C.c_int),
        "petstrip_run": ([vp, fp, C.c_int, C.c_int, OUTPUT_CB, PROGRESS_CB, vp],
                         C.c_int),
        "petstrip_sensitivity": ([vp, C.c_int, fp], C.c_int),
        "petstrip_stats_get": ([vp, C.POINTER(CStats)], C.c_int),
        "petstrip_footprints": ([vp, fp, fp], C.c_int),
        "petstrip_forward": ([vp, fp], C.c_int),
        "petstrip_backward": ([vp, fp, fp, C.POINTER(C.c_double),
                               C.POINTER(C.c_double)], C.c_int),
        "petstrip_set_sigma": ([vp, dp, dp, C.c_int], C.c_int),
        "petstrip_phantom_events": ([C.POINTER(CScanner), fp, i32, C.c_uint64,
                                     i64, fp, C.c_int], C.c_int),
        "petstrip_phantom_events_device": ([C.POINTER(CScanner), fp, i32, C.c_uint64,
                                            i64, vp, C.c_int], C.c_int),
        "petstrip_set_phantom_events": ([vp, fp, i32, C.c_uint64, i64], C.c_int),
        "petstrip_get_events": ([vp, fp], C.c_int),
        "petstrip_phantom_generate": ([C.POINTER(CScanner), fp, i32, i32, C.c_uint64, i64,
                                       C.POINTER(CPhantomGrid), C.POINTER(CPhantomOutput),
                                       C.c_int], C.c_int),
        "petstrip_phantom_tan_kernel_map": ([C.POINTER(CScanner), C.POINTER(CPhantomGrid),
                                             dp, dp, dp, fp, C.c_int], C.c_int),
        "petstrip_read_text_responses": ([C.c_char_p, i32, fp, fp, i64,
                                          C.POINTER(C.c_int64)], C.c_int),
        "petstrip_parse_text_floats": ([C.c_char_p, i32, fp, C.POINTER(C.c_int32)],
                                       C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise PetstripError(rc, _lib.petstrip_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def read_text_responses(path, columns=3, with_lors=False):
    """Responses of a text file as the reference's fast text reader reads
    them (src/util/text_parser.h:29-175, src/2d/strip/reconstruction.h:
    169-196): float32 (n, 3) of (z_u, z_d, dl); columns = 5 for
    "i j z_u z_d dl" lines, whose (i, j) come back as int16 (n, 2) with
    with_lors.  Host only."""
    lib = load_library()
    n = C.c_int64(0)
    _check(lib.petstrip_read_text_responses(os.fsencode(path), columns, None, None, 0,
                                            C.byref(n)))
    ev = np.zeros((n.value, 3), np.float32)
    lors = np.zeros((n.value, 2), np.int16) if with_lors else None
    got = C.c_int64(0)
    _check(lib.petstrip_read_text_responses(os.fsencode(path), columns, _ptr(ev), _ptr(lors),
                                            n.value, C.byref(got)))
    ev, lors = ev[:got.value], (lors[:got.value] if with_lors else None)
    return (ev, lors) if with_lors else ev


def parse_text_floats(text, n=1):
    """n space-separated floats of a string, read as the reference's fast
    text reader reads them (text_parser.h:140-186); returns (values,
    characters consumed)."""
    lib = load_library()
    out = np.zeros(n, np.float32)
    used = C.c_int32(0)
    raw = text.encode() if isinstance(text, str) else bytes(text)
    _check(lib.petstrip_parse_text_floats(raw, n, _ptr(out), C.byref(used)))
    return out, used.value


class Scanner:
    """Strip scanner geometry (src/2d/strip/scanner.h:39-67)."""

    def __init__(self, radius, scintillator_length, n_y_pixels, n_z_pixels,
                 pixel_height, pixel_width, sigma_z, sigma_dl, center_y=0.0,
                 center_z=0.0):
        self.radius = float(radius)
        self.scintillator_length = float(scintillator_length)
        self.n_y_pixels = int(n_y_pixels)
        self.n_z_pixels = int(n_z_pixels)
        self.pixel_height = float(pixel_height)
        self.pixel_width = float(pixel_width)
        self.sigma_z = float(sigma_z)
        self.sigma_dl = float(sigma_dl)
        self.center_y = float(center_y)
        self.center_z = float(center_z)

    @classmethod
    def square(cls, radius, length, n_pixels, pixel_size, sigma_z, sigma_dl):
        return cls(radius, length, n_pixels, n_pixels, pixel_size, pixel_size,
                   sigma_z, sigma_dl)

    @property
    def total_n_pixels(self):
        return self.n_y_pixels * self.n_z_pixels

    def c(self):
        return CScanner(self.radius, self.scintillator_length, self.n_y_pixels,
                        self.n_z_pixels, self.pixel_height, self.pixel_width,
                        self.sigma_z, self.sigma_dl, self.center_y, self.center_z)


def device_count():
    lib = load_library()
    c = C.c_int(0)
    _check(lib.petstrip_device_count(C.byref(c)))
    return c.value


PROBE_LIB_PATH = os.path.join(HERE, "libpetstrip_probe.so")
PROBE_GLOBAL_F32, PROBE_GLOBAL_F64, PROBE_GLOBAL_U64, PROBE_LDS_U64 = 0, 1, 2, 3
PROBE_LDS_U32, PROBE_LDS_F32 = 4, 5
_probe_lib = None


def load_probe_library(path=PROBE_LIB_PATH):
    """libpetstrip_probe.so (include/petstrip_probe.h)."""
    global _probe_lib
    if _probe_lib is not None:
        return _probe_lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(path)
    lib.petstrip_probe_atomics.argtypes = [C.c_int, C.c_int, C.c_int32, C.c_int32,
                                           C.c_int32, C.POINTER(C.c_double),
                                           C.POINTER(C.c_double)]
    lib.petstrip_probe_atomics.restype = C.c_int
    _probe_lib = lib
    return lib


def probe_atomics(kind, width=512, height=512, reps=1024, device=0):
    """Device-wide rate of one kind of atomic update (PROBE_*), in atomics
    per second, and the timed launch's milliseconds."""
    lib = load_probe_library()
    rate, ms = C.c_double(0.0), C.c_double(0.0)
    rc = lib.petstrip_probe_atomics(int(device), int(kind), int(width), int(height),
                                    int(reps), C.byref(rate), C.byref(ms))
    if rc != 0:
        raise PetstripError(rc, f"atomic probe failed with HIP error {rc}")
    return rate.value, ms.value


def phantom_events(scanner, regions, n, seed=1, n_threads=None):
    """Seeded detected events (n x 3 float32: z_u, z_d, dl).

    regions: iterable of (x, y, a, b, angle_deg, intensity); earlier regions
    lie on top (non-additive), as in src/2d/geometry/phantom.h:193-203.
    """
    lib = load_library()
    reg = np.ascontiguousarray(np.asarray(regions, dtype=np.float64).reshape(-1, 6))
    out = np.empty((int(n), 3), dtype=np.float32)
    if n_threads is None:
        n_threads = min(os.cpu_count() or 1, 16)
    sc = scanner.c()
    _check(lib.petstrip_phantom_events(C.byref(sc), _ptr(reg), reg.shape[0],
                                       C.c_uint64(seed), C.c_int64(int(n)),
                                       _ptr(out), int(n_threads)))
    return out


REGION_ELLIPSE, REGION_RECTANGLE, REGION_POINT = 0, 1, 2
PHANTOM_ADDITIVE, PHANTOM_DETECTED = 1, 2


def phantom_grid(n_columns, n_rows, pixel_size, origin=(0.0, 0.0), tan_bins=0,
                 max_tan=0.0):
    """PixelGrid centred on origin (lower-left corner at
    -pixel_size * (n_columns, n_rows) / 2 + origin), row 0 at the bottom."""
    return CPhantomGrid(int(n_columns), int(n_rows), float(pixel_size),
                        -pixel_size * n_columns / 2 + origin[0],
                        -pixel_size * n_rows / 2 + origin[1], int(tan_bins), float(max_tan))


def phantom_generate(scanner, regions, n, seed=1, additive=False, detected=False,
                     grid=None, full=False, images=False, device=-1):
    """The phantom command's generator (src/2d/strip/phantom_cmd.cpp:136-444).

    regions: (kind, x, y, a, b, angle_deg, intensity) with kind REGION_*.
    detected: n counts detected events (else emissions).  With grid and
    images, also the emitted / detected / naive images (n_rows x n_columns,
    row 0 at the bottom) and, with grid.tan_bins > 0, the tan-bins map.
    device < 0: host threads.  Returns a dict."""
    lib = load_library()
    reg = np.ascontiguousarray(np.asarray(regions, dtype=np.float64).reshape(-1, 7))
    n = int(n)
    res = {"responses": np.empty((n, 3), np.float32)}
    if full:
        res["exact"] = np.empty((n, 3), np.float32)
        res["events"] = np.empty((n, 4), np.float32)
    if images:
        shape = (grid.n_rows, grid.n_columns)
        for k in ("emitted", "detected", "naive"):
            res[k] = np.empty(shape, np.int32)
        if grid.tan_bins > 0:
            res["tan_bins_map"] = np.empty((grid.tan_bins,) + shape, np.int32)
    out = CPhantomOutput(*[_ptr(res.get(k)) for k in
                           ("responses", "exact", "events", "emitted", "detected", "naive",
                            "tan_bins_map")], 0, 0)
    flags = (PHANTOM_ADDITIVE if additive else 0) | (PHANTOM_DETECTED if detected else 0)
    sc = scanner.c()
    _check(lib.petstrip_phantom_generate(C.byref(sc), _ptr(reg), reg.shape[0], flags,
                                         C.c_uint64(seed), C.c_int64(n),
                                         C.byref(grid) if grid is not None else None,
                                         C.byref(out), int(device)))
    m = out.n_responses
    res["responses"] = res["responses"][:m]
    if full:
        res["exact"] = res["exact"][:m]
        res["events"] = res["events"][:m]
    res["n_emitted"] = out.n_emitted
    return res


def phantom_tan_kernel_map(scanner, grid, kernel_point=(0.0, 0.0), scale=1.0, device=0):
    """The --tan-bins kernel map (tan_bins x n_rows x n_columns float32)."""
    lib = load_library()
    out = np.empty((grid.tan_bins, grid.n_rows, grid.n_columns), np.float32)
    sc = scanner.c()
    _check(lib.petstrip_phantom_tan_kernel_map(C.byref(sc), C.byref(grid),
                                               float(kernel_point[0]),
                                               float(kernel_point[1]), float(scale),
                                               _ptr(out), int(device)))
    return out


class Reconstruction:
    """List-mode MLEM on one GPU (src/2d/strip/reconstruction.h:32-154).

    rho starts at 1; responses are (z_u, z_d, dl) float32 triples.
    """

    def __init__(self, scanner, device=0, accumulation=ACCUM_FP32,
                 update=UPDATE_CPU, path=PATH_AUTO, row_scan=False, force_bands=False,
                 kernel=KERNEL_GAUSSIAN, row_table=True, horner=True):
        self.lib = load_library()
        self.scanner = scanner
        self._h = C.c_void_p()
        opts = COptions(accumulation, update, path, 1 if row_scan else 0,
                        int(force_bands), kernel, 0 if row_table else 1,
                        0 if horner else 1)
        sc = scanner.c()
        _check(self.lib.petstrip_create(C.byref(sc), device, C.byref(opts),
                                        C.byref(self._h)))
        self.n_events = 0

    def close(self):
        if self._h:
            self.lib.petstrip_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def shape(self):
        return (self.scanner.n_y_pixels, self.scanner.n_z_pixels)

    def device_name(self):
        buf = C.create_string_buffer(256)
        _check(self.lib.petstrip_device_name(self._h, buf, 256))
        return buf.value.decode()

    def stream(self):
        s = C.c_void_p()
        _check(self.lib.petstrip_stream(self._h, C.byref(s)))
        return s.value

    def set_options(self, accumulation=ACCUM_FP32, update=UPDATE_CPU, path=PATH_AUTO,
                    row_scan=False, force_bands=False, kernel=KERNEL_GAUSSIAN,
                    row_table=True, horner=True):
        o = COptions(accumulation, update, path, 1 if row_scan else 0,
                     int(force_bands), kernel, 0 if row_table else 1, 0 if horner else 1)
        _check(self.lib.petstrip_set_options(self._h, C.byref(o)))

    def set_events(self, responses):
        r = np.ascontiguousarray(responses, dtype=np.float32).reshape(-1, 3)
        _check(self.lib.petstrip_set_events(self._h, _ptr(r), C.c_int64(r.shape[0])))
        self.n_events = r.shape[0]

    def set_phantom_events(self, regions, n, seed=1):
        """Generates n seeded phantom events on the device (the generator of
        phantom_events, one lane per slot) and prepares them as set_events."""
        reg = np.ascontiguousarray(np.asarray(regions, dtype=np.float64).reshape(-1, 6))
        _check(self.lib.petstrip_set_phantom_events(self._h, _ptr(reg), reg.shape[0],
                                                    C.c_uint64(seed), C.c_int64(int(n))))
        self.n_events = int(n)

    def events(self):
        """The responses on the device, n x 3 float32."""
        out = np.empty((self.n_events, 3), dtype=np.float32)
        _check(self.lib.petstrip_get_events(self._h, _ptr(out)))
        return out

    def set_events_device(self, ptr, n):
        _check(self.lib.petstrip_set_events_device(self._h, C.c_void_p(ptr),
                                                   C.c_int64(int(n))))
        self.n_events = int(n)

    @property
    def rho(self):
        out = np.empty(self.shape, dtype=np.float32)
        _check(self.lib.petstrip_get_rho(self._h, _ptr(out)))
        return out

    @rho.setter
    def rho(self, value):
        v = np.ascontiguousarray(value, dtype=np.float32).reshape(self.shape)
        _check(self.lib.petstrip_set_rho(self._h, _ptr(v)))

    def __call__(self, n_iterations=1):
        """Runs n_iterations MLEM iterations (reconstruction.h:108-154)."""
        _check(self.lib.petstrip_iterate(self._h, int(n_iterations)))
        return self

    iterate = __call__

    def run(self, rho, n_blocks, n_iterations_in_block, output=None, progress=None):
        """The reference GPU run(): output(iteration, image) at the reference's
        output iterations (-1 = sensitivity); progress(completed, finished)."""
        rho = np.ascontiguousarray(rho, dtype=np.float32).reshape(self.shape)
        W, H = self.scanner.n_z_pixels, self.scanner.n_y_pixels

        def _out(user, it, img, w, h):
            if output is None:
                return 0
            a = np.ctypeslib.as_array(img, shape=(h, w)).copy()
            r = output(it, a)
            return 1 if r else 0

        def _prog(user, completed, finished):
            if progress is not None:
                progress(completed, bool(finished))

        ocb, pcb = OUTPUT_CB(_out), PROGRESS_CB(_prog)
        _check(self.lib.petstrip_run(self._h, _ptr(rho), int(n_blocks),
                                     int(n_iterations_in_block), ocb, pcb, None))
        return rho

    def sensitivity(self, mode=0):
        out = np.empty(self.shape, dtype=np.float32)
        _check(self.lib.petstrip_sensitivity(self._h, int(mode), _ptr(out)))
        return out

    def stats(self):
        s = CStats()
        _check(self.lib.petstrip_stats_get(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in CStats._fields_}

    def footprints(self):
        nb = np.empty(self.n_events, dtype=np.int32)
        ne = np.empty(self.n_events, dtype=np.int32)
        _check(self.lib.petstrip_footprints(self._h, _ptr(nb), _ptr(ne)))
        return nb, ne

    def forward(self):
        lam = np.empty(self.n_events, dtype=np.float32)
        _check(self.lib.petstrip_forward(self._h, _ptr(lam)))
        return lam

    def backward(self, g=None, want_rho=True, want_sigma=True, want_loglik=False):
        gr = np.empty(self.shape, dtype=np.float32) if want_rho else None
        gs = (C.c_double * 2)() if want_sigma else None
        ll = C.c_double(0.0) if want_loglik else None
        garr = None if g is None else np.ascontiguousarray(g, dtype=np.float32)
        _check(self.lib.petstrip_backward(
            self._h, _ptr(garr), _ptr(gr),
            gs if gs is not None else None,
            C.byref(ll) if ll is not None else None))
        return (gr, (gs[0], gs[1]) if gs is not None else None,
                ll.value if ll is not None else None)

    def set_sigma(self, sigma_z, sigma_dl, freeze_support=False):
        _check(self.lib.petstrip_set_sigma(self._h, float(sigma_z), float(sigma_dl),
                                           1 if freeze_support else 0))
