"""Synthetic inputs of the 3D hybrid reconstruction: a barrel's LOR geometry
in the layout of the reference's GeometrySOA
(src/2d/barrel/geometry_soa.h:22-48,178-199) and seeded (LOR, z_up, z_dn,
dl) responses of a cylindrical phantom.

The reference builds its geometry files with 2d_barrel_geometry, which needs
Boost.Geometry (src/2d/barrel/geometry_cmd.cpp:185-240); this module gives a
geometry of the same form for tests and measurement: one ring of detectors,
each LOR a segment between two detector centres (the second detector's centre
first, as the reference's GeometrySOA does, geometry_soa.h:137-139), with the
grid pixels whose centres lie within a band around the LOR and inside the
field of view, ordered by their position t along the segment (what the
reconstruction's binary searches need, reconstruction.h:129-145).
"""
import numpy as np


class HybridGeometry:
    """GeometrySOA arrays plus the voxel grid.

    segments: (n_lors, 4) float32 start x, y, end x, y; widths (n_lors,)
    float32; info_begin, info_end (n_lors,) int64; pixels (n, 2) int16 (x, y);
    positions (n,) float32; weights (n,) float32 (recomputed by the
    reconstruction unless it is told to keep them).  LOR index of (first,
    second), first >= second: first (first + 1) / 2 + second.
    """

    def __init__(self, n_detectors, n_columns, n_rows, pixel_size, lower_left, z_left,
                 n_planes, segments, widths, info_begin, info_end, pixels, positions,
                 weights, ring_radius):
        self.n_detectors = int(n_detectors)
        self.n_columns, self.n_rows = int(n_columns), int(n_rows)
        self.pixel_size = float(pixel_size)
        self.lower_left = (float(lower_left[0]), float(lower_left[1]))
        self.z_left, self.n_planes = float(z_left), int(n_planes)
        self.segments = np.ascontiguousarray(segments, dtype=np.float32)
        self.widths = np.ascontiguousarray(widths, dtype=np.float32)
        self.info_begin = np.ascontiguousarray(info_begin, dtype=np.int64)
        self.info_end = np.ascontiguousarray(info_end, dtype=np.int64)
        self.pixels = np.ascontiguousarray(pixels, dtype=np.int16)
        self.positions = np.ascontiguousarray(positions, dtype=np.float32)
        self.weights = np.ascontiguousarray(weights, dtype=np.float32)
        self.ring_radius = float(ring_radius)

    @property
    def n_lors(self):
        return self.n_detectors * (self.n_detectors + 1) // 2

    @property
    def n_pixel_infos(self):
        return len(self.positions)

    @property
    def n_voxels(self):
        return self.n_columns * self.n_rows * self.n_planes


def lor_index(first, second):
    first = np.asarray(first, dtype=np.int64)
    return first * (first + 1) // 2 + np.asarray(second, dtype=np.int64)


def detector_centres(n_detectors, ring_radius):
    """Centres of a ring of detectors, detector 0 at -y, counter-clockwise
    (the angles of the reference's ring builder, scanner_builder.h:76-80)."""
    a = 2 * np.pi * np.arange(n_detectors) / n_detectors - np.pi / 2
    return np.stack([ring_radius * np.cos(a), ring_radius * np.sin(a)], axis=1)


def ring_geometry(n_detectors=64, ring_radius=0.25, detector_width=0.007, n_pixels=64,
                  pixel_size=0.005, n_planes=None, fov_radius=None):
    """One ring of n_detectors, an n_pixels x n_pixels grid of pixel_size
    centred on the axis, n_planes planes (default n_pixels) centred in z.
    A pixel belongs to a LOR when its centre lies within width/2 + half a
    pixel diagonal of the LOR's line, inside the field of view and between
    the two detectors."""
    n_planes = n_pixels if n_planes is None else n_planes
    half = n_pixels * pixel_size / 2
    fov = half if fov_radius is None else fov_radius
    centres = detector_centres(n_detectors, ring_radius)
    idx = np.arange(n_pixels)
    cx, cy = np.meshgrid(-half + (idx + 0.5) * pixel_size, -half + (idx + 0.5) * pixel_size)
    px, py = np.meshgrid(idx, idx)
    inside = cx * cx + cy * cy <= fov * fov
    cx, cy, px, py = cx[inside], cy[inside], px[inside], py[inside]
    band = detector_width / 2 + pixel_size / np.sqrt(2)
    n_lors = n_detectors * (n_detectors + 1) // 2
    segments = np.zeros((n_lors, 4), np.float32)
    widths = np.full(n_lors, detector_width, np.float32)
    begin = np.zeros(n_lors, np.int64)
    end = np.zeros(n_lors, np.int64)
    pix, pos = [], []
    offset = 0
    for first in range(n_detectors):
        for second in range(first + 1):
            li = first * (first + 1) // 2 + second
            s, e = centres[second], centres[first]
            if first == second:
                # a self-LOR has no pixels; the last one gets a chord through
                # the centre (to the opposite detector) and one pixel info,
                # so that a range reaching one past the previous LOR's end
                # (reconstruction.h:133-138) stays inside the arrays
                e = centres[(first + n_detectors // 2) % n_detectors]
            segments[li] = (s[0], s[1], e[0], e[1])
            begin[li] = offset
            if first != second:
                d = e - s
                length = np.hypot(d[0], d[1])
                u = d / length
                dist = np.abs((cx - s[0]) * (-u[1]) + (cy - s[1]) * u[0])
                t = ((cx - s[0]) * u[0] + (cy - s[1]) * u[1]) / length
                sel = (dist <= band) & (t >= 0) & (t <= 1)
                if np.any(sel):
                    t32 = t[sel].astype(np.float32)
                    order = np.argsort(t32, kind="stable")
                    pix.append(np.stack([px[sel][order], py[sel][order]], axis=1))
                    pos.append(t32[order])
                    offset += int(sel.sum())
            elif first == n_detectors - 1:
                c = n_pixels // 2
                pix.append(np.array([[c, c]]))
                pos.append(np.array([0.5], np.float32))
                offset += 1
            end[li] = offset
    pixels = np.concatenate(pix).astype(np.int16)
    positions = np.concatenate(pos).astype(np.float32)
    return HybridGeometry(n_detectors, n_pixels, n_pixels, pixel_size, (-half, -half),
                          -n_planes * pixel_size / 2, n_planes, segments, widths, begin, end,
                          pixels, positions, np.ones(len(positions), np.float32), ring_radius)


def cylinder_responses(geometry, n, sigma_z=0.01, sigma_dl=0.04, length=0.3, radius=None,
                       half_z=None, seed=1, max_tan=0.5):
    """n responses (lors int16 (n, 2) first > second; zzd float32 (n, 3):
    z_up, z_dn, dl) of emissions uniform in a cylinder (radius default half
    the grid, half length default a third of the grid's z extent), the
    in-plane direction uniform, the axial slope tan uniform in +-max_tan.  The
    LOR is the detector pair nearest the line's two ring crossings; z_up
    belongs to the LOR's first (higher) detector, the segment's end; errors
    N(0, sigma_z) and N(0, sigma_dl) are added and responses with |z| beyond
    length / 2 are redrawn."""
    g = geometry
    rng = np.random.default_rng(seed)
    ext = g.n_columns * g.pixel_size / 2
    radius = ext / 2 if radius is None else radius
    half_z = g.n_planes * g.pixel_size / 6 if half_z is None else half_z
    centres = detector_centres(g.n_detectors, g.ring_radius)
    out_l = np.zeros((0, 2), np.int16)
    out_z = np.zeros((0, 3), np.float32)
    while len(out_l) < n:
        m = 2 * (n - len(out_l)) + 64
        r = radius * np.sqrt(rng.random(m))
        a = 2 * np.pi * rng.random(m)
        x, y = r * np.cos(a), r * np.sin(a)
        z = rng.uniform(-half_z, half_z, m)
        phi = np.pi * rng.random(m)
        tan = rng.uniform(-max_tan, max_tan, m)
        ux, uy = np.cos(phi), np.sin(phi)
        # crossings of the line with the ring: p + s u, |p + s u| = R
        b = x * ux + y * uy
        c = x * x + y * y - g.ring_radius ** 2
        root = np.sqrt(b * b - c)
        step = 2 * np.pi / g.n_detectors
        dets = []
        for sgn in (1, -1):
            qx, qy = x + (-b + sgn * root) * ux, y + (-b + sgn * root) * uy
            ang = np.arctan2(qy, qx) + np.pi / 2
            dets.append(np.round(ang / step).astype(np.int64) % g.n_detectors)
        d1, d2 = dets
        first, second = np.maximum(d1, d2), np.minimum(d1, d2)
        ok = first != second
        li = lor_index(first, second)
        ok &= g.info_end[li] > g.info_begin[li]
        s, e = centres[second], centres[first]
        dv = e - s
        seg_len = np.hypot(dv[:, 0], dv[:, 1])
        u = dv / seg_len[:, None]
        mid = (s + e) / 2
        up = (x - mid[:, 0]) * u[:, 0] + (y - mid[:, 1]) * u[:, 1]
        R = seg_len / 2
        sec = np.sqrt(1 + tan * tan)
        z_up = z + (R - up) * tan + rng.normal(0, sigma_z, m)
        z_dn = z - (R + up) * tan + rng.normal(0, sigma_z, m)
        dl = -2 * up * sec + rng.normal(0, sigma_dl, m)
        ok &= (np.abs(z_up) < length / 2) & (np.abs(z_dn) < length / 2)
        lors = np.stack([first, second], axis=1)[ok].astype(np.int16)
        zzd = np.stack([z_up, z_dn, dl], axis=1)[ok].astype(np.float32)
        out_l = np.concatenate([out_l, lors])
        out_z = np.concatenate([out_z, zzd])
    return np.ascontiguousarray(out_l[:n]), np.ascontiguousarray(out_z[:n])


# ---------------------------------------------------------------------------
# The reference's binary geometry file (2d_barrel_geometry's output, read by
# 3d_hybrid_reconstruction --geometry): native little-endian fields,
#   uint32 magic "PETg" (float geometry), int16 n_detectors,
#   grid: int16 n_columns, int16 n_rows, float pixel_size, float lower-left x, y
#   then per LOR, in index order: int16 first, int16 second, float start x, y,
#   end x, y, float width, uint64 n, n x (int16 x, y, float t, distance,
#   fill, weight)
# (src/2d/barrel/geometry.h:55-104, lor_geometry.h:46-88,
#  2d/geometry/pixel_grid.h:63-77, line_segment.h:33-56, util/bstream.h).
GEOMETRY_MAGIC = b"PETg"
_PIXEL_INFO = np.dtype([("x", "<i2"), ("y", "<i2"), ("t", "<f4"), ("distance", "<f4"),
                        ("fill", "<f4"), ("weight", "<f4")])


def read_geometry(path, n_planes=None, z_left=None):
    """A HybridGeometry from the reference's geometry file; the voxel grid's
    planes default as 3d_hybrid_reconstruction's do (n_planes = the larger
    pixel count, z_left = -n_planes pixel_size / 2; reconstruction_cmd.cpp:
    287-315)."""
    data = open(path, "rb").read()
    if data[:4] != GEOMETRY_MAGIC:
        raise ValueError("unknown binary geometry file")
    n_det = int(np.frombuffer(data, "<i2", 1, 4)[0])
    n_col, n_row = (int(v) for v in np.frombuffer(data, "<i2", 2, 6))
    ps, llx, lly = (float(v) for v in np.frombuffer(data, "<f4", 3, 10))
    n_lors = n_det * (n_det + 1) // 2
    segments = np.zeros((n_lors, 4), np.float32)
    widths = np.zeros(n_lors, np.float32)
    begin = np.zeros(n_lors, np.int64)
    end = np.zeros(n_lors, np.int64)
    infos = []
    off, total = 22, 0
    head = np.dtype([("first", "<i2"), ("second", "<i2"), ("seg", "<f4", 4),
                     ("width", "<f4"), ("n", "<u8")])
    while off + head.itemsize <= len(data):
        h = np.frombuffer(data, head, 1, off)[0]
        off += head.itemsize
        n = int(h["n"])
        first, second = int(h["first"]), int(h["second"])
        li = first * (first + 1) // 2 + second
        if not (0 <= second <= first < n_det):
            raise ValueError("geometry file: LOR out of range")
        segments[li] = h["seg"]
        widths[li] = h["width"]
        pi = np.frombuffer(data, _PIXEL_INFO, n, off) if n else np.zeros(0, _PIXEL_INFO)
        off += n * _PIXEL_INFO.itemsize
        infos.append((li, pi))
    # pixel infos laid out in LOR index order (the GeometrySOA layout)
    infos.sort(key=lambda x: x[0])
    pix, pos, wts = [], [], []
    for li, pi in infos:
        begin[li] = total
        total += len(pi)
        end[li] = total
        pix.append(np.stack([pi["x"], pi["y"]], axis=1))
        pos.append(pi["t"])
        wts.append(pi["weight"])
    # LORs absent from the file keep an empty range at the running offset
    last = 0
    for li in range(n_lors):
        if end[li] == 0 and begin[li] == 0:
            begin[li] = end[li] = last
        last = end[li]
    n_planes = max(n_col, n_row) if n_planes is None else int(n_planes)
    z_left = -n_planes * ps / 2 if z_left is None else float(z_left)
    cat = (lambda xs, t, shape: np.concatenate(xs).astype(t) if xs else np.zeros(shape, t))
    return HybridGeometry(n_det, n_col, n_row, ps, (llx, lly), z_left, n_planes, segments,
                          widths, begin, end, cat(pix, np.int16, (0, 2)),
                          cat(pos, np.float32, (0,)), cat(wts, np.float32, (0,)), 0.0)


def write_geometry(g, path):
    """The geometry in the reference's binary format (every LOR, in index
    order; distance and fill written as 0)."""
    with open(path, "wb") as f:
        f.write(GEOMETRY_MAGIC)
        f.write(np.array([g.n_detectors, g.n_columns, g.n_rows], "<i2").tobytes())
        f.write(np.array([g.pixel_size, g.lower_left[0], g.lower_left[1]], "<f4").tobytes())
        for first in range(g.n_detectors):
            for second in range(first + 1):
                li = first * (first + 1) // 2 + second
                b, e = int(g.info_begin[li]), int(g.info_end[li])
                f.write(np.array([first, second], "<i2").tobytes())
                f.write(np.asarray(g.segments[li], "<f4").tobytes())
                f.write(np.array([g.widths[li]], "<f4").tobytes())
                f.write(np.array([e - b], "<u8").tobytes())
                pi = np.zeros(e - b, _PIXEL_INFO)
                pi["x"], pi["y"] = g.pixels[b:e, 0], g.pixels[b:e, 1]
                pi["t"] = g.positions[b:e]
                pi["weight"] = g.weights[b:e]
                f.write(pi.tobytes())


# --- system matrices and scanner description files (the --system inputs) ----

def write_system_matrix(path, n_pixels_in_row, n_emissions, n_detectors, elements,
                        n_tof_positions=None):
    """Writes a 2D barrel system matrix file in the full form (magic "PETP",
    or "TOFP" with n_tof_positions; src/2d/barrel/sparse_matrix.h:43-67): the
    header (with the TOF position count after the detectors), then per group
    of consecutive elements with the same (a, b) detectors: a, b (uint16),
    the count (uint32) and per element [the TOF position (uint32)], x, y
    (uint16) and hits (uint32).  elements: rows (a, b, x, y, hits), or
    (a, b, position, x, y, hits) with TOF, written in the given order."""
    tof = n_tof_positions is not None
    el = np.asarray(elements, dtype=np.int64).reshape(-1, 6 if tof else 5)
    if not tof:
        el = np.insert(el, 2, 0, axis=1)
    with open(path, "wb") as f:
        f.write(b"TOFP" if tof else b"PETP")
        f.write(np.array([n_pixels_in_row, n_emissions, n_detectors], "<u4").tobytes())
        if tof:
            f.write(np.array([n_tof_positions], "<u4").tobytes())
        fields = ([("position", "<u4")] if tof else []) + [("x", "<u2"), ("y", "<u2"),
                                                           ("hits", "<u4")]
        k = 0
        while k < len(el):
            j = k
            while j < len(el) and el[j, 0] == el[k, 0] and el[j, 1] == el[k, 1]:
                j += 1
            f.write(np.array([el[k, 0], el[k, 1]], "<u2").tobytes())
            f.write(np.array([j - k], "<u4").tobytes())
            rec = np.zeros(j - k, dtype=fields)
            if tof:
                rec["position"] = el[k:j, 2]
            rec["x"], rec["y"], rec["hits"] = el[k:j, 3], el[k:j, 4], el[k:j, 5]
            f.write(rec.tobytes())
            k = j


def synthetic_system_matrix(g, n_emissions=10000, seed=1, plane=0):
    """Elements (a, b, x, y, hits) of a system matrix for geometry g: every
    pixel info of every LOR with hits falling off with its distance to the
    LOR (and with `plane`, for per-plane matrices), plus a few pixels the
    geometry's LOR lacks, the LORs in shuffled order, each LOR's pixels
    shuffled, detectors of half the LORs swapped (a < b)."""
    rng = np.random.default_rng(seed)
    rows = []
    for first in range(g.n_detectors):
        for second in range(first):
            li = first * (first + 1) // 2 + second
            b, e = int(g.info_begin[li]), int(g.info_end[li])
            if e <= b:
                continue
            px = g.pixels[b:e].astype(np.int64)
            t = g.positions[b:e].astype(np.float64)
            hits = (n_emissions * 0.05 * np.exp(-4 * (t - 0.5) ** 2) / (1 + plane)
                    * rng.uniform(0.5, 1.5, e - b)).astype(np.int64) + 1
            a_, b_ = (first, second) if rng.random() < 0.5 else (second, first)
            lor_rows = [(a_, b_, int(x), int(y), int(h)) for (x, y), h in zip(px, hits)]
            if rng.random() < 0.2:  # a pixel outside the geometry's range
                lor_rows.append((a_, b_, int(rng.integers(0, g.n_columns)),
                                 int(rng.integers(0, g.n_rows)), int(rng.integers(1, 50))))
            rng.shuffle(lor_rows)
            rows.append(lor_rows)
    rng.shuffle(rows)
    return np.array([r for lr in rows for r in lr], dtype=np.int64)


def synthetic_tof_matrix(g, n_emissions=10000, seed=1, n_tof_positions=3):
    """Elements (a, b, position, x, y, hits) of a TOF system matrix: those of
    synthetic_system_matrix with each (LOR, pixel) split over one to
    n_tof_positions TOF positions, so that a LOR holds its pixels several
    times."""
    rng = np.random.default_rng(seed + 7919)
    rows = []
    for a, b, x, y, h in synthetic_system_matrix(g, n_emissions, seed):
        k = int(rng.integers(1, n_tof_positions + 1))
        pos = rng.choice(n_tof_positions, size=k, replace=False)
        for p in pos:
            rows.append((a, b, int(p), x, y, max(1, int(h * rng.uniform(0.2, 1.0)))))
    return np.array(rows, dtype=np.int64)


def write_detector_file(path, n_detectors, ring_radius, width=0.007, height=0.019):
    """A scanner description file (ScannerBuilder::deserialize,
    src/2d/barrel/scanner_builder.h:21-35): per detector one line of its
    four corners "x y x y x y x y", a width x height rectangle centred on
    the ring's detector centre, its height along the radius."""
    c = detector_centres(n_detectors, ring_radius)
    lines = []
    for x, y in c:
        r = np.hypot(x, y)
        ux, uy = x / r, y / r  # radial
        tx, ty = -uy, ux       # tangential
        pts = []
        for sw, sh in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
            pts += [x + sw * width / 2 * tx + sh * height / 2 * ux,
                    y + sw * width / 2 * ty + sh * height / 2 * uy]
        lines.append(" ".join(f"{v:.9g}" for v in pts))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
