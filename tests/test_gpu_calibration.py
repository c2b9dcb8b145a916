"""Sigma calibration (alternating MLEM rounds and gradient steps on
(sigma_z, sigma_dl)) on the 160 x 160 / 4 mm scanner: every accepted step
and every MLEM round raises the log-likelihood, and the widths move from the
start (1.5x) toward the simulated ones."""
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
