"""The 3D hybrid command's --system paths on the CPU: the library's system
matrix reader, per-plane weight loader and geometry-from-matrix builder
(csrc/system_matrix.cpp) and the oracle's multi-plane iteration
(oracle/hybrid_oracle.c), bit for bit against what the reference's own code
made of the same files (tests/golden/hybrid_system.npz, written by
tests/golden/make_hybrid_system_golden.py from oracle/_ref)."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as ol

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
SYS = os.path.join(GOLD, "hybrid_system")
sys.path.insert(0, GOLD)
import make_hybrid_system_golden as ms  # noqa: E402

F = np.load(os.path.join(GOLD, "hybrid_system.npz"))
MATS = [os.path.join(SYS, f"matrix_{p}.bin") for p in range(ms.N_MATRICES)]
DET = os.path.join(SYS, "detectors.txt")
TOF = os.path.join(SYS, "matrix_tof.bin")


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def frames_matrix(fr):
    return np.stack([fr["first"], fr["second"], fr["up"], fr["right"], fr["tan"], fr["sec"],
                     fr["begin"], fr["end"], fr["plane_begin"], fr["plane_end"]],
                    axis=1).astype(np.float64)


@pytest.fixture(scope="module")
def case_inputs(pkg):
    g, lors, zzd = ms.inputs(pkg)
    assert np.array_equal(lors, F["lors"]) and same_bits(zzd, F["zzd"])
    return g, lors, zzd


def raw_plane_weights(pkg, g, n):
    return np.stack([pkg.hybrid.SystemMatrix(p).weights(g) for p in MATS[:n]])


def matrix_geometry(pkg, inactive=(), path=MATS[0]):
    m = pkg.hybrid.SystemMatrix(path)
    centres = pkg.hybrid.read_detector_centres(DET)
    return m.geometry(centres, 0.01, n_planes=ms.N_PLANES, z_left=-0.06, inactive=inactive)


def test_matrix_file_header(pkg):
    m = pkg.hybrid.SystemMatrix(MATS[0])
    assert (m.n_pixels_in_row, m.n_detectors, m.n_emissions) == (24, 24, 20000)
    assert m.n_elements > 4000
    assert pkg.hybrid.read_detector_centres(DET).shape == (24, 2)


def test_triangular_and_bad_files_refused(pkg, tmp_path):
    p = tmp_path / "tri.bin"
    p.write_bytes(b"PETp" + np.array([12, 100, 24], "<u4").tobytes())
    with pytest.raises(pkg.PetstripError, match="full form"):
        pkg.hybrid.SystemMatrix(p)
    p.write_bytes(b"XXXX" + np.array([12, 100, 24], "<u4").tobytes())
    with pytest.raises(pkg.PetstripError, match="invalid file type"):
        pkg.hybrid.SystemMatrix(p)
    with pytest.raises(pkg.PetstripError, match="cannot open"):
        pkg.hybrid.SystemMatrix(tmp_path / "none.bin")


@pytest.mark.parametrize("name,n", [("am", ms.N_MATRICES), ("a1", 1)])
def test_plane_weights_vs_reference(pkg, case_inputs, name, n):
    """load_weights_from_matrix_file for each weight plane (the library),
    then the sensitivity and normalisation (the oracle), against the
    reference's normalised weights and sensitivity."""
    g, _, _ = case_inputs
    raw = raw_plane_weights(pkg, g, n)
    # the matrices hold pixels the geometry's LORs lack: those are dropped
    assert np.all(raw >= 0) and np.count_nonzero(raw) < sum(
        pkg.hybrid.SystemMatrix(p).n_elements for p in MATS[:n])
    w, s = ol.hybrid_plane_weights(g, raw, 0)
    assert same_bits(w, F[f"{name}_weights"])
    assert same_bits(s.reshape(F[f"{name}_sensitivity"].shape), F[f"{name}_sensitivity"])


@pytest.mark.parametrize("name,inactive", [("b", ()), ("bi", (3, 10)), ("bt", ())])
def test_geometry_from_matrix_vs_reference(pkg, name, inactive):
    """The geometry of a matrix and a scanner description file (the
    library) against the reference's GeometrySOA(matrix, centres, grid,
    inactive): segments, pixel-info ranges, pixels and positions bit for
    bit; the LORs of inactive detectors are empty; a TOF matrix repeats a
    LOR's pixels once per TOF position."""
    h = matrix_geometry(pkg, inactive, TOF if name == "bt" else MATS[0])
    ni = h.n_pixel_infos
    assert same_bits(h.segments, F[f"{name}_segments"])
    assert np.array_equal(h.info_begin, F[f"{name}_begin"])
    assert np.array_equal(h.info_end, F[f"{name}_end"])
    assert np.array_equal(h.pixels, F[f"{name}_pixels"][:ni])
    assert same_bits(h.positions, F[f"{name}_positions"][:ni])
    ref_w = F[f"{name}_weights"][0]
    if name == "bi":
        assert ni < F["bi_n_pixel_infos"]  # the reference counts skipped elements too
        for d in inactive:
            for other in range(24):
                a, b = max(d, other), min(d, other)
                li = a * (a + 1) // 2 + b
                assert h.info_end[li] == h.info_begin[li]
        assert same_bits(h.weights, ref_w[:ni])  # sensitivity one
    else:
        assert ni == F[f"{name}_n_pixel_infos"]
        w, s = ol.hybrid_plane_weights(h, h.weights[None, :], 0)
        assert same_bits(w[0], ref_w)
        assert same_bits(s.reshape(F[f"{name}_sensitivity"].shape), F[f"{name}_sensitivity"])
    if name == "bt":
        repeats = sum(int(e - b) - len({tuple(p) for p in h.pixels[b:e]})
                      for b, e in zip(h.info_begin, h.info_end))
        assert repeats > ni // 3


@pytest.mark.parametrize("name", ["am", "a1", "b", "bi", "bt"])
def test_oracle_iteration_vs_reference(pkg, case_inputs, name):
    """The oracle's frames and iteration with the reference's weights and
    sensitivity: the frames and both images bit for bit."""
    g, lors, zzd = case_inputs
    if name.startswith("b"):
        g = matrix_geometry(pkg, (3, 10) if name == "bi" else (),
                            TOF if name == "bt" else MATS[0])
    w = F[f"{name}_weights"][:, :g.n_pixel_infos]
    s = F[f"{name}_sensitivity"]
    fr = ol.hybrid_frames(g, ms.SZ, ms.SDL, w[0], lors, zzd)
    assert np.array_equal(frames_matrix(fr), F[f"{name}_frames"])
    assert fr["plane_begin"].min() >= 1  # plane 0 lies outside the reference's arrays
    if w.shape[0] > 1:
        rho = ol.hybrid_mlem_planes(g, ms.SZ, ms.SDL, ms.LENGTH, w, s, fr, ms.N_ITER)
    else:
        rho = ol.hybrid_mlem(g, ms.SZ, ms.SDL, ms.LENGTH, w[0], s[0], fr, ms.N_ITER)
    assert same_bits(rho, F[f"{name}_rho"])
    if name == "am":
        # the weight planes matter: their order reversed gives another image
        bad = ol.hybrid_mlem_planes(g, ms.SZ, ms.SDL, ms.LENGTH, w[::-1], s, fr, ms.N_ITER)
        assert np.linalg.norm(bad[-1] - rho[-1]) > 1e-3 * np.linalg.norm(rho[-1])
        # FP64 twin within FP32 rounding
        r64 = ol.hybrid_mlem_planes(g, ms.SZ, ms.SDL, ms.LENGTH, w, s, fr, ms.N_ITER, f64=True)
        assert np.linalg.norm(r64[-1] - rho[-1]) < 1e-5 * np.linalg.norm(r64[-1])
