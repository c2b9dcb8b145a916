"""Sigma calibration (alternating MLEM rounds and gradient steps on
(sigma_z, sigma_dl)): every accepted step and every MLEM round raises the
log-likelihood and the widths move from the start (1.5x) toward the
simulated ones; and, with the image held at the phantom's expected image of
detected emissions, the estimator recovers the simulated widths."""
import math

import pytest

pytestmark = pytest.mark.gpu


def test_calibration_raises_loglik_and_moves_toward_truth(gpu_pkg):
    pkg = gpu_pkg
    name = "strip-160px-4mm-1e6ev"
    _, _, _, _, phantom, seed = pkg.WORKLOADS[name]
    sc = pkg.workload_scanner(name)
    print(f"case: {name}, 200000 events, 6 rounds", flush=True)
    with pkg.Reconstruction(sc) as rec:
        rec.set_phantom_events(phantom, 200000, seed=seed)
        res = pkg.calibrate_sigma(rec, (0.015, 0.06), rounds=6)
        assert (rec.stats()["n_events"], res["mlem_iterations"]) == (200000, 30)
    h = res["history"]
    seq = []
    for e in h:
        seq += [e["loglik"], e["loglik_after_step"]]
        print(e, flush=True)
    for a, b in zip(seq, seq[1:]):
        assert b >= a - 1e-9 * abs(a)
    for k, s0, st in ((0, 0.015, sc.sigma_z), (1, 0.06, sc.sigma_dl)):
        fin = (res["sigma_z"], res["sigma_dl"])[k]
        assert abs(math.log(fin / st)) < 0.75 * abs(math.log(s0 / st))


def test_calibration_on_expected_image_recovers_sigma(gpu_pkg):
    """With rho fixed at the expected image of detected emissions (the
    phantom's intensity times the detection probability, rasterised), the
    gradient ascent in sigma alone converges to the widths the events were
    simulated with: (0.01, 0.04) within 5 %, from (0.015, 0.06), on 2e6
    events at 512 x 512 pixels of 1 mm.  (Alternating with MLEM instead lets
    the sharpening image absorb part of the blur.)"""
    pkg = gpu_pkg
    name = "strip-512px-1mm-2e7ev"
    _, _, _, _, phantom, seed = pkg.WORKLOADS[name]
    sc = pkg.workload_scanner(name)
    n = 2_000_000
    print(f"case: {name}, {n} events, expected image, 20 rounds of sigma steps", flush=True)
    with pkg.Reconstruction(sc) as rec:
        rec.set_phantom_events(phantom, n, seed=seed)
        rec.rho = pkg.expected_image(sc, phantom, n, rec.sensitivity(0))
        res = pkg.calibrate_sigma(rec, (0.015, 0.06), rounds=20, mlem_per_round=0)
    for e in res["history"]:
        print(f"  round {e['round']}: ({e['sigma_z']:.5f}, {e['sigma_dl']:.5f}) "
              f"loglik {e['loglik']:.6e}", flush=True)
    seq = []
    for e in res["history"]:
        seq += [e["loglik"], e["loglik_after_step"]]
    assert all(b >= a for a, b in zip(seq, seq[1:]))
    assert abs(res["sigma_z"] / sc.sigma_z - 1) < 0.05
    assert abs(res["sigma_dl"] / sc.sigma_dl - 1) < 0.05
