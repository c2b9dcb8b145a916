"""The 3D hybrid reconstruction on the GPU (include/petstrip_hybrid.h) against
the oracle (oracle/hybrid_oracle.c, bitwise equal to the reference's
PET3D::Hybrid::Reconstruction) and the reference fixtures.

Bit-exact: the LOR-pixel weights and sensitivity (host, the reference's
float expressions) and the frame events (device, the same float expressions
without contraction: pixel ranges and plane ranges are the reference's
discrete decisions).  Within tolerance: the images, whose kernel uses the
device's exp and atan and whose sums run in another order (relative L2
<= 1e-5 against the FP32 oracle, <= 1e-4 against the FP64 twin)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_hybrid_golden as mg  # noqa: E402


def rel(a, b):
    return float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b))


def frames_equal(a, b):
    keys = ["first", "second", "up", "right", "tan", "sec", "begin", "end", "plane_begin",
            "plane_end"]
    return len(a) == len(b) and all(np.array_equal(a[k], b[k]) for k in keys)


@pytest.mark.parametrize("name", sorted(mg.CASES))
def test_hybrid_vs_reference_fixture(gpu_pkg, name):
    import oracle_lib as ol
    pkg = gpu_pkg
    c = mg.CASES[name]
    sz, sdl, length, n_iter, sens_one = c[5], c[6], c[7], c[9], c[11]
    g, lors, zzd = mg.build(name)
    f = np.load(os.path.join(HERE, "golden", name + ".npz"))
    print(f"case {name}: {len(lors)} responses, {g.n_pixel_infos} pixel infos", flush=True)
    flags = pkg.hybrid.SENSITIVITY_ONE if sens_one else 0
    with pkg.HybridReconstruction(g, sz, sdl, length, flags=flags) as rec:
        w, s = rec.weights()
        assert np.array_equal(w.view(np.uint32), f["weights"].view(np.uint32))
        assert np.array_equal(s.view(np.uint32), f["sensitivity"].view(np.uint32))
        rec.set_responses(lors, zzd)
        ev = rec.events()
        want = ol.hybrid_frames(g, sz, sdl, w, lors, zzd)
        assert frames_equal(ev, want)
        for it in range(n_iter):
            rec(1)
            r = rel(rec.rho, f["rho"][it])
            print(f"  iteration {it + 1}: rel L2 vs reference {r:.2e}", flush=True)
            assert r < 1e-5


def test_hybrid_3d_sensitivity_vs_oracle(gpu_pkg):
    """An external 3D sensitivity (the command's --sensitivity): weights,
    frames and two iterations against the oracle, which matches the
    reference bit for bit on this case (test_hybrid_oracle.py); a 2D map set
    afterwards replaces the volume again."""
    import oracle_lib as ol
    pkg = gpu_pkg
    hg = pkg.hybrid_geometry
    g = hg.ring_geometry(n_detectors=40, ring_radius=0.21, detector_width=0.008, n_pixels=28,
                         pixel_size=0.009, n_planes=48)
    lors, zzd = hg.cylinder_responses(g, 1200, sigma_z=0.011, sigma_dl=0.045, length=0.28,
                                      seed=17)
    vol = np.random.default_rng(17).uniform(0.5, 1.5, (24, 28, 28)).astype(np.float32)
    print("case: 28x28x48 voxels, 24-plane sensitivity volume, 1200 responses", flush=True)
    w_ref, _ = ol.hybrid_weights(g, external=vol[0])
    fr = ol.hybrid_frames(g, 0.011, 0.045, w_ref, lors, zzd)
    want = ol.hybrid_mlem(g, 0.011, 0.045, 0.28, w_ref, vol, fr, 2)
    with pkg.HybridReconstruction(g, 0.011, 0.045, 0.28) as rec:
        w2d, s2d = rec.weights()
        with pytest.raises(Exception):
            rec.set_sensitivity_3d(vol[:23])
        rec.set_sensitivity_3d(vol)
        w, s = rec.weights()
        assert np.array_equal(w.view(np.uint32), w_ref.view(np.uint32))
        assert np.all(s == 1.0)
        rec.set_responses(lors, zzd)
        assert frames_equal(rec.events(), fr)
        for it in range(2):
            rec(1)
            r = rel(rec.rho, want[it])
            print(f"  iteration {it + 1}: rel L2 vs oracle {r:.2e}", flush=True)
            assert r < 1e-5
        # back to the plain 2D sensitivity: the oracle's 2D run
        rec.set_weights(w2d, s2d)
        rec.rho = np.ones(rec.shape, np.float32)
        rec(1)
        want2d = ol.hybrid_mlem(g, 0.011, 0.045, 0.28, w2d, s2d, fr, 1)
        assert rel(rec.rho, want2d[0]) < 1e-5


def test_hybrid_vs_oracle_f64(gpu_pkg):
    import oracle_lib as ol
    pkg = gpu_pkg
    hg = pkg.hybrid_geometry
    g = hg.ring_geometry(n_detectors=64, ring_radius=0.25, detector_width=0.007, n_pixels=48,
                         pixel_size=0.005)
    lors, zzd = hg.cylinder_responses(g, 40000, seed=11)
    sz, sdl, length = 0.01, 0.04, 0.3
    print(f"case: 48^3 voxels, 64 detectors, 4e4 responses, 3 iterations", flush=True)
    with pkg.HybridReconstruction(g, sz, sdl, length) as rec:
        rec.set_responses(lors, zzd)
        st = rec.stats()
        w, s = rec.weights()
        fr = rec.events()
        gpu = []
        for _ in range(3):
            rec(1)
            gpu.append(rec.rho)
    assert st["n_events"] == len(fr) > 39000 and st["n_voxels"] > 0
    o32 = ol.hybrid_mlem(g, sz, sdl, length, w, s, fr, 3)
    w64, s64 = ol.hybrid_weights(g, f64=True)
    o64 = ol.hybrid_mlem(g, sz, sdl, length, w64, s64, fr, 3, f64=True)
    for it in range(3):
        r32, r64 = rel(gpu[it], o32[it]), rel(gpu[it], o64[it])
        print(f"  iteration {it + 1}: rel L2 vs oracle f32 {r32:.2e}, vs f64 {r64:.2e}",
              flush=True)
        assert r32 < 1e-5 and r64 < 1e-4


def test_hybrid_long_ranges(gpu_pkg):
    """Wide detectors: ranges of more than 192 pixel infos (the kernel then
    forms the per-pixel terms at every voxel instead of in its LDS table) and
    events of more than 1024 voxels (whose later voxels the backprojection
    recomputes instead of reading them from the LDS cache)."""
    import oracle_lib as ol
    pkg = gpu_pkg
    hg = pkg.hybrid_geometry
    g = hg.ring_geometry(n_detectors=32, ring_radius=0.2, detector_width=0.018, n_pixels=48,
                         pixel_size=0.004, n_planes=40)
    lors, zzd = hg.cylinder_responses(g, 3000, sigma_z=0.012, sigma_dl=0.05, length=0.3,
                                      seed=21)
    print("case: 32 detectors of 18 mm, 48x48x40 voxels of 4 mm, 3000 responses", flush=True)
    with pkg.HybridReconstruction(g, 0.012, 0.05, 0.3) as rec:
        rec.set_responses(lors, zzd)
        fr = rec.events()
        w, s = rec.weights()
        gpu = []
        for _ in range(2):
            rec(1)
            gpu.append(rec.rho)
    n_pix = np.minimum(fr["end"], g.n_pixel_infos) - fr["begin"]
    n_pl = fr["plane_end"] - fr["plane_begin"]
    assert 0.2 < np.mean(n_pix > 192) < 0.9 and np.mean(n_pix * n_pl > 1024) > 0.9
    o32 = ol.hybrid_mlem(g, 0.012, 0.05, 0.3, w, s, fr, 2)
    for it in range(2):
        r = rel(gpu[it], o32[it])
        print(f"  iteration {it + 1}: rel L2 vs oracle f32 {r:.2e}", flush=True)
        assert r < 1e-5


def test_hybrid_events_and_run_schedule(gpu_pkg):
    """Frame events handed over by the caller give the image of the
    responses; run() follows the reference's block schedule; with sensitivity
    one every event adds one to the image."""
    pkg = gpu_pkg
    hg = pkg.hybrid_geometry
    g = hg.ring_geometry(n_detectors=48, ring_radius=0.22, n_pixels=32, pixel_size=0.0075)
    lors, zzd = hg.cylinder_responses(g, 5000, seed=4)
    print("case: events handed over, run schedule, sensitivity one", flush=True)
    with pkg.HybridReconstruction(g, 0.01, 0.04, 0.3,
                                  flags=pkg.hybrid.SENSITIVITY_ONE) as rec:
        rec.set_responses(lors, zzd)
        ev = rec.events()
        rec(2)
        a = rec.rho
        n_ev = rec.stats()["n_events"]
        rec.rho = np.ones(rec.shape, np.float32)
        rec.set_events(ev)
        assert frames_equal(rec.events(), ev)
        rec(2)
        b = rec.rho
        rec.rho = np.ones(rec.shape, np.float32)
        rec(4)
        four = rec.rho
        outs = []
        img = rec.run(2, 2, output=lambda it, im: outs.append((it, im.shape)))
    assert rel(b, a) < 1e-6
    assert abs(a.sum(dtype=np.float64) / n_ev - 1) < 1e-4
    assert [o[0] for o in outs] == [1, 2, 4]
    assert all(o[1] == (32, 32, 32) for o in outs)
    assert rel(img, four) < 1e-6
    assert np.all(np.isfinite(img))


def test_hybrid_rejects_bad_input(gpu_pkg):
    pkg = gpu_pkg
    hg = pkg.hybrid_geometry
    g = hg.ring_geometry(n_detectors=16, ring_radius=0.2, n_pixels=16, pixel_size=0.01)
    with pkg.HybridReconstruction(g, 0.01, 0.04, 0.3) as rec:
        with pytest.raises(pkg.PetstripError):
            rec.set_responses(np.array([[2, 5]], np.int16), np.zeros((1, 3), np.float32))
        with pytest.raises(pkg.PetstripError):
            rec.set_responses(np.array([[16, 5]], np.int16), np.zeros((1, 3), np.float32))
    bad = hg.ring_geometry(n_detectors=16, ring_radius=0.2, n_pixels=16, pixel_size=0.01)
    bad.pixels[0, 0] = 16
    with pytest.raises(pkg.PetstripError):
        pkg.HybridReconstruction(bad, 0.01, 0.04, 0.3)
