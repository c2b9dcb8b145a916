"""The 3D hybrid oracle (oracle/hybrid_oracle.c) against the reference's own
reconstruction (PET3D::Hybrid::Reconstruction, src/3d/hybrid/
reconstruction.h): bitwise on the committed fixtures (tests/golden/hybrid_*.npz,
written by tests/golden/make_hybrid_golden.py from oracle/_ref), and, where
the harness oracle/_ref/libref_hybrid.so was built, on a case the fixtures
do not hold.  Plus the invariants the mathematics fixes and mistakes that
must fail."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as ol

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_hybrid_golden as mg  # noqa: E402


def load(name):
    return np.load(os.path.join(HERE, "golden", name + ".npz"))


def frames_matrix(fr):
    return np.stack([fr["first"], fr["second"], fr["up"], fr["right"], fr["tan"], fr["sec"],
                     fr["begin"], fr["end"], fr["plane_begin"], fr["plane_end"]],
                    axis=1).astype(np.float64)


def oracle_run(g, c, lors, zzd):
    sz, sdl, length, n_iter, sens_one = c[5], c[6], c[7], c[9], c[11]
    w, s = ol.hybrid_weights(g, sensitivity_one=sens_one)
    fr = ol.hybrid_frames(g, sz, sdl, w, lors, zzd)
    rho = ol.hybrid_mlem(g, sz, sdl, length, w, s, fr, n_iter)
    return w, s, fr, rho


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("name", sorted(mg.CASES))
def test_oracle_bitwise_vs_reference_fixture(name):
    g, lors, zzd = mg.build(name)
    f = load(name)
    assert np.array_equal(lors, f["lors"]) and same_bits(zzd, f["zzd"])
    w, s, fr, rho = oracle_run(g, mg.CASES[name], lors, zzd)
    assert same_bits(w, f["weights"])
    assert same_bits(s, f["sensitivity"])
    assert np.array_equal(frames_matrix(fr), f["frames"])
    for it in range(rho.shape[0]):
        assert same_bits(rho[it], f["rho"][it]), f"iteration {it + 1}"
    # some events reach one past their LOR's last pixel info
    li = fr["first"].astype(np.int64) * (fr["first"] + 1) // 2 + fr["second"]
    assert np.any(fr["end"] > g.info_end[li])


@pytest.mark.skipif(not ol.have_ref_hybrid(), reason="oracle/_ref/libref_hybrid.so not built")
def test_oracle_bitwise_vs_reference_fresh_case(pkg):
    hg = pkg.hybrid_geometry
    g = hg.ring_geometry(n_detectors=40, ring_radius=0.21, detector_width=0.008, n_pixels=28,
                         pixel_size=0.009, n_planes=20)
    lors, zzd = hg.cylinder_responses(g, 1200, sigma_z=0.011, sigma_dl=0.045, length=0.28,
                                      seed=99)
    r = ol.ref_hybrid(g, 0.011, 0.045, 0.28, lors, zzd, 2)
    w, s = ol.hybrid_weights(g)
    fr = ol.hybrid_frames(g, 0.011, 0.045, w, lors, zzd)
    rho = ol.hybrid_mlem(g, 0.011, 0.045, 0.28, w, s, fr, 2)
    assert same_bits(w, r["weights"]) and same_bits(s, r["sensitivity"])
    assert np.array_equal(frames_matrix(fr), r["frames"])
    assert same_bits(rho, r["rho"])


def _case_3d_sensitivity(pkg):
    """40 detectors, 28 x 28 pixels of 9 mm x 48 planes, an external volume
    of 24 planes; the planes reach past the cylinder's kernels on both sides,
    so no event touches plane 0 (which would read one plane past the
    reference's volume)."""
    hg = pkg.hybrid_geometry
    g = hg.ring_geometry(n_detectors=40, ring_radius=0.21, detector_width=0.008, n_pixels=28,
                         pixel_size=0.009, n_planes=48)
    lors, zzd = hg.cylinder_responses(g, 1200, sigma_z=0.011, sigma_dl=0.045, length=0.28,
                                      seed=17)
    vol = np.random.default_rng(17).uniform(0.5, 1.5, (24, 28, 28)).astype(np.float32)
    return g, lors, zzd, vol


@pytest.mark.skipif(not ol.have_ref_hybrid(), reason="oracle/_ref/libref_hybrid.so not built")
def test_oracle_bitwise_vs_reference_3d_sensitivity(pkg):
    """--sensitivity (reconstruction_cmd.cpp:505-519): weights normalised by
    the volume's plane 0, denominators over plane |iz - 24|."""
    g, lors, zzd, vol = _case_3d_sensitivity(pkg)
    r = ol.ref_hybrid(g, 0.011, 0.045, 0.28, lors, zzd, 2, sensitivity_3d=vol)
    w, s = ol.hybrid_weights(g, external=vol[0])
    fr = ol.hybrid_frames(g, 0.011, 0.045, w, lors, zzd)
    assert fr["plane_begin"].min() >= 1
    rho = ol.hybrid_mlem(g, 0.011, 0.045, 0.28, w, vol, fr, 2)
    assert same_bits(w, r["weights"]) and same_bits(s, vol[0])
    assert same_bits(r["sensitivity"], vol[0])
    assert np.array_equal(frames_matrix(fr), r["frames"])
    assert same_bits(rho, r["rho"])
    # the planes beyond 0 matter: the volume's plane 0 repeated gives another image
    flat = ol.hybrid_mlem(g, 0.011, 0.045, 0.28, w, np.repeat(vol[:1], 24, axis=0), fr, 1)
    assert not np.allclose(flat[0], rho[0], rtol=1e-3)


def test_sensitivity_one_conserves_events():
    """With sensitivity one the denominator is the sum of the weights, so
    every event with a positive denominator adds exactly one to the image
    (up to the reference's repeated-pixel rule and float rounding)."""
    name = "hybrid_24px_sens1"
    g, lors, zzd = mg.build(name)
    w, s, fr, rho = oracle_run(g, mg.CASES[name], lors, zzd)
    assert np.all(s == 1)
    for it in range(rho.shape[0]):
        assert abs(rho[it].sum(dtype=np.float64) / len(fr) - 1) < 1e-4
    assert np.all(rho >= 0)


def test_f64_twin_close_to_f32():
    name = "hybrid_32px_48det"
    c = mg.CASES[name]
    g, lors, zzd = mg.build(name)
    w, s, fr, rho = oracle_run(g, c, lors, zzd)
    w64, s64 = ol.hybrid_weights(g, f64=True)
    # the float distance to the LOR is a difference of numbers near the ring
    # radius, so the weights far out in the tail carry ~1e-5 relative error
    assert np.linalg.norm(w64 - w) / np.linalg.norm(w64) < 1e-5
    rho64 = ol.hybrid_mlem(g, c[5], c[6], c[7], w64, s64, fr, c[9], f64=True)
    for it in range(c[9]):
        rel = np.linalg.norm(rho64[it] - rho[it]) / np.linalg.norm(rho64[it])
        assert rel < 1e-4, (it, rel)


def test_planted_mistakes_fail():
    """A segment taken from the other detector, or a kernel width off by a
    tenth of a percent, no longer reproduces the reference."""
    name = "hybrid_32px_48det"
    c = mg.CASES[name]
    f = load(name)
    g, lors, zzd = mg.build(name)
    g.segments = np.ascontiguousarray(g.segments[:, [2, 3, 0, 1]])
    w, s = ol.hybrid_weights(g)
    fr = ol.hybrid_frames(g, c[5], c[6], w, lors, zzd)
    rho = ol.hybrid_mlem(g, c[5], c[6], c[7], w, s, fr, 1)
    assert not same_bits(rho[0], f["rho"][0])
    g, lors, zzd = mg.build(name)
    w, s = ol.hybrid_weights(g)
    fr = ol.hybrid_frames(g, c[5], c[6], w, lors, zzd)
    rho = ol.hybrid_mlem(g, c[5] * 1.001, c[6], c[7], w, s, fr, 1)
    assert not same_bits(rho[0], f["rho"][0])


def test_geometry_file_round_trip_and_reference_reader(pkg, tmp_path):
    """write_geometry writes the reference's binary geometry format:
    read_geometry reads it back unchanged, and the reference's own reader
    (Geometry(ibstream) into a GeometrySOA) sees the same arrays."""
    hg = pkg.hybrid_geometry
    g = hg.ring_geometry(n_detectors=24, ring_radius=0.2, n_pixels=20, pixel_size=0.012,
                         n_planes=14)
    g.weights = np.random.default_rng(3).random(g.n_pixel_infos).astype(np.float32)
    path = str(tmp_path / "ring24.bin")
    hg.write_geometry(g, path)
    h = hg.read_geometry(path, n_planes=14, z_left=g.z_left)
    assert (h.n_detectors, h.n_columns, h.n_rows, h.n_planes) == (24, 20, 20, 14)
    assert h.pixel_size == np.float32(g.pixel_size)
    for k in ["segments", "widths", "info_begin", "info_end", "pixels", "positions",
              "weights"]:
        assert np.array_equal(getattr(h, k), getattr(g, k)), k
    if not ol.have_ref_hybrid():
        pytest.skip("oracle/_ref/libref_hybrid.so not built")
    r = ol.ref_hybrid_read_geometry(path)
    assert (r["n_detectors"], r["n_columns"], r["n_rows"]) == (24, 20, 20)
    assert np.array_equal(r["grid"], np.array([g.pixel_size, *g.lower_left], np.float32))
    for k in ["segments", "widths", "info_begin", "info_end", "pixels", "positions",
              "weights"]:
        assert np.array_equal(r[k], getattr(g, k)), k
