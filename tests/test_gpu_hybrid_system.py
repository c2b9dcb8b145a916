"""The 3D hybrid command's --system paths on the GPU: per-plane weights from
system matrices (petstrip_hybrid_set_plane_weights) and the geometry of a
matrix and a scanner description file, against what the reference's own
code made of the same files (tests/golden/hybrid_system.npz, written by
tests/golden/make_hybrid_system_golden.py).

Bit-exact: the normalised weights and sensitivity (host, the reference's
float expressions) and the frame events.  Within tolerance: the images
(relative L2 <= 1e-5: the device's exp and atan and another summation
order)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
SYS = os.path.join(GOLD, "hybrid_system")
sys.path.insert(0, GOLD)
import make_hybrid_system_golden as ms  # noqa: E402

MATS = [os.path.join(SYS, f"matrix_{p}.bin") for p in range(ms.N_MATRICES)]
DET = os.path.join(SYS, "detectors.txt")
TOF = os.path.join(SYS, "matrix_tof.bin")


def rel(a, b):
    return float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b))


def frames_matrix(ev):
    return np.stack([ev["first"], ev["second"], ev["up"], ev["right"], ev["tan"], ev["sec"],
                     ev["begin"], ev["end"], ev["plane_begin"], ev["plane_end"]],
                    axis=1).astype(np.float64)


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


CASES = {  # name: (matrices, geometry from the matrix, inactive, sensitivity mode)
    "am": (ms.N_MATRICES, False, (), 0),
    "a1": (1, False, (), 0),
    "b": (1, True, (), 0),
    "bi": (1, True, (3, 10), 1),
    # a TOF matrix: each LOR holds its pixels up to three times, and every
    # entry of a pixel backprojects the weight of its last entry
    "bt": (1, True, (), 0),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_system_paths_vs_reference(gpu_pkg, name):
    pkg = gpu_pkg
    F = np.load(os.path.join(GOLD, "hybrid_system.npz"))
    n_mat, from_matrix, inactive, mode = CASES[name]
    g, lors, zzd = ms.inputs(pkg)
    if from_matrix:
        g = pkg.hybrid.SystemMatrix(TOF if name == "bt" else MATS[0]).geometry(
            pkg.hybrid.read_detector_centres(DET), 0.01, n_planes=ms.N_PLANES,
            z_left=-0.06, inactive=inactive)
        raw = g.weights[None, :]
    else:
        raw = np.stack([pkg.hybrid.SystemMatrix(p).weights(g) for p in MATS[:n_mat]])
    ni = g.n_pixel_infos
    print(f"case {name}: {raw.shape[0]} weight planes, {ni} pixel infos, {len(lors)} "
          f"responses", flush=True)
    with pkg.HybridReconstruction(g, ms.SZ, ms.SDL, ms.LENGTH) as rec:
        rec.set_plane_weights(raw, sensitivity_mode=mode)
        w, s = rec.plane_weights()
        assert same_bits(w, F[f"{name}_weights"][:, :ni])
        assert same_bits(s, F[f"{name}_sensitivity"])
        assert same_bits(rec.weights()[0], w[0])  # the frame filter's plane
        rec.set_responses(lors, zzd)
        assert np.array_equal(frames_matrix(rec.events()), F[f"{name}_frames"])
        for it in range(ms.N_ITER):
            rec(1)
            r = rel(rec.rho, F[f"{name}_rho"][it])
            print(f"  iteration {it + 1}: rel L2 vs reference {r:.2e}", flush=True)
            assert r < 1e-5


def test_plane_weights_order_and_errors(gpu_pkg):
    """The weight planes are not interchangeable (reversed, another image);
    a plane count other than one or n_planes / 2, or a sensitivity volume
    missing in mode 2, is refused; one plane afterwards gives the one-set
    image again."""
    import oracle_lib as ol
    pkg = gpu_pkg
    F = np.load(os.path.join(GOLD, "hybrid_system.npz"))
    g, lors, zzd = ms.inputs(pkg)
    raw = np.stack([pkg.hybrid.SystemMatrix(p).weights(g) for p in MATS])
    with pkg.HybridReconstruction(g, ms.SZ, ms.SDL, ms.LENGTH) as rec:
        with pytest.raises(pkg.PetstripError):
            rec.set_plane_weights(raw[:2])
        with pytest.raises(pkg.PetstripError):
            rec.set_plane_weights(raw, sensitivity_mode=2)
        rec.set_plane_weights(raw[::-1])
        rec.set_responses(lors, zzd)
        rec(1)
        assert rel(rec.rho, F["am_rho"][0]) > 1e-3
        # mode 2: the reference's sensitivity volume given back gives its image
        rec.set_plane_weights(raw, sensitivity_mode=2, sensitivity=F["am_sensitivity"])
        rec.rho = np.ones(rec.shape, np.float32)
        rec(1)
        assert rel(rec.rho, F["am_rho"][0]) < 1e-5
        assert same_bits(rec.plane_weights()[1], F["am_sensitivity"])
        rec.set_plane_weights(raw[:1])
        w, s = rec.weights()
        assert same_bits(w, F["a1_weights"][0])
        assert same_bits(rec.plane_weights()[0], F["a1_weights"])
        rec.rho = np.ones(rec.shape, np.float32)
        rec(1)
        fr = ol.hybrid_frames(g, ms.SZ, ms.SDL, w, lors, zzd)
        want = ol.hybrid_mlem(g, ms.SZ, ms.SDL, ms.LENGTH, w, s, fr, 1)
        assert rel(rec.rho, want[0]) < 1e-5
