"""Detector-resolution calibration by alternating MLEM and gradient ascent.

The strip kernel's widths (sigma_z, sigma_dl) enter the model through the
3-sigma support, the normalisation and the exponent of every kernel value
(src/2d/strip/gaussian_kernel.h:19-79).  Given list-mode events, the Poisson
list-mode log-likelihood

    L(rho, sigma) = sum_e log lambda_e(rho, sigma) - sum_p rho_p

is raised in turn in the image (MLEM iterations, which increase L at fixed
sigma) and in sigma (one gradient step at fixed rho, the derivative taken with
the support held at the current sigma).  Steps are taken in log sigma:
u = log sigma, dL/du = sigma dL/dsigma, divided by the number of events so
that a step is O(1) far from the optimum; for a Gaussian width the per-event
curvature of L in u is about -2, so a step of dL/du / (2N) is close to a
Newton step.  Each step is checked at once with a forward pass at the new
sigma: a step that lowers L is halved, up to three times, and dropped if L
still falls.
"""
import math
import time


def calibrate_sigma(rec, sigma0, rounds=20, mlem_per_round=5, rate=0.5, max_step=0.1,
                    support="start", grow=2.0, log=None):
    """Runs `rounds` rounds of `mlem_per_round` MLEM iterations followed by a
    step on (sigma_z, sigma_dl), starting from sigma0, on the events and image
    already in `rec` (a Reconstruction).  The step is rate x sigma dL/dsigma /
    N in log sigma, clipped to max_step; after an accepted full step the rate
    grows by `grow` (the curvature of L in log sigma is much weaker than a
    single Gaussian's, since the image spreads each event's likelihood), and
    every halving of a rejected step halves it.  Returns a dict with the final
    sigmas, the per-round history and timings; rec holds the final image and
    sigma on return."""
    if support not in ("start", "current"):
        raise ValueError("support must be 'start' or 'current'")
    n = max(int(rec.n_events), 1)
    sz, sdl = float(sigma0[0]), float(sigma0[1])
    frozen = support == "start"
    rec.set_sigma(sz, sdl)
    history = []
    t_mlem = t_grad = 0.0
    t_start = time.perf_counter()
    for r in range(rounds):
        t0 = time.perf_counter()
        if mlem_per_round > 0:
            rec(mlem_per_round)
        t1 = time.perf_counter()
        _, g, ll = rec.backward(want_rho=False, want_sigma=True, want_loglik=True)
        t_grad += time.perf_counter() - t1
        t_mlem += t1 - t0
        du = [max(-max_step, min(max_step, rate * s * gi / n))
              for s, gi in zip((sz, sdl), g)]
        accepted, halvings = False, 0
        for halvings in range(4):
            nz, ndl = sz * math.exp(du[0]), sdl * math.exp(du[1])
            rec.set_sigma(nz, ndl, freeze_support=frozen)
            ll_new = rec.backward(want_rho=False, want_sigma=False, want_loglik=True)[2]
            if ll_new >= ll:
                accepted = True
                break
            du = [d / 2 for d in du]
        if not accepted:
            nz, ndl = sz, sdl
            rec.set_sigma(sz, sdl, freeze_support=frozen)
            ll_new = ll
        rate = rate * grow if accepted and halvings == 0 else rate / 2 ** (halvings + 1 - accepted)
        entry = {"round": r + 1, "sigma_z": sz, "sigma_dl": sdl, "loglik": ll,
                 "grad_sigma": [float(g[0]), float(g[1])], "next_sigma_z": nz,
                 "next_sigma_dl": ndl, "loglik_after_step": ll_new, "rate": rate}
        history.append(entry)
        if log:
            log(entry)
        sz, sdl = nz, ndl
    return {"sigma_z": sz, "sigma_dl": sdl, "history": history,
            "seconds": time.perf_counter() - t_start, "seconds_mlem": t_mlem,
            "seconds_gradient": t_grad, "mlem_iterations": rounds * mlem_per_round}
