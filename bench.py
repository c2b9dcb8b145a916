"""Benchmark: list-mode MLEM events/s/iteration on one MI355X.

Workload (default): strip scanner R = 0.43 m, L = 0.5 m, sigma_z = 0.01 m,
sigma_dl = 0.04 m; image 512 x 512 pixels of 1 mm; 2e7 seeded events from a
six-ellipse phantom; one step = one MLEM iteration over all events with the
events and the image resident in device memory.

  python bench.py [--steps K] [--warmup W] [--workload NAME] [--gpus 1]

Prints one JSON line.  value = events * K / (time of K iterations).
roofline: the fused iteration kernel is FP32-VALU bound (no contraction to
put on MFMA, image and events are a few hundred MB against ~5e12 flop);
achieved = algorithmic flop per launch (sum_e 15 N_box + 41 N_ell + 50) /
mean launch time measured with HIP events on the library's stream.
cpu_baseline: the reference CPU algorithm, the reference's own sources
compiled by oracle/Makefile (or, with --cpu-kind port, the oracle's
restatement), on a bounded prefix of the same events: iterations 1-3 from
rho = 1, each timed, events/s/iter from the median of iterations 2-3; with
the OpenMP threads used, the CPUs in the process's affinity mask, nproc and
the CPU model.
parity: the GPU image over the same prefix against that CPU image, and over a
shorter prefix against FP64 runs of the same algorithm (relative L2 and the
largest per-pixel relative error, per iteration).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("MLEM events/s/iter at 512², 2e7 ev vs host CPU; grad ms; "
          "rel L2 err vs FP64")
VALU_PEAK_TFLOPS = 157.3   # MI355X FP32 vector peak
HBM_PEAK_GBS = 8000.0


def atomic_report(pkg, st, n_pix, kernel_ms, accumulation="fp32"):
    """Atomic updates of one MLEM iteration against the device's measured
    atomic rates.  Pass 2 makes one LDS atomic per ellipse pixel (ds_add_u32
    in the FP32 iteration's 32-bit windows, ds_add_u64 in the other modes);
    the flush at most one global atomic per window pixel (zero pixels are
    skipped; FP64 in the FP32 and compensated iterations).  The reference's
    GPU path makes one global atomic per ellipse pixel
    (src/2d/strip/cuda/warp_granularity.cuh:145-150), priced at the FP32
    rate.  Rates are over the whole kernel time, so each fraction is the
    share of the kernel that the updates would take at the probe's rate."""
    lds_kind = "u32" if accumulation == "fp32" else "u64"
    flush_kind = "u64" if accumulation == "fixed64" else "f64"
    try:
        g32, _ = pkg.probe_atomics(pkg.PROBE_GLOBAL_F32, n_pix, n_pix, reps=1024)
        g_rate, _ = pkg.probe_atomics(
            pkg.PROBE_GLOBAL_U64 if flush_kind == "u64" else pkg.PROBE_GLOBAL_F64,
            n_pix, n_pix, reps=1024)
        l_rate, _ = pkg.probe_atomics(
            pkg.PROBE_LDS_U32 if lds_kind == "u32" else pkg.PROBE_LDS_U64, reps=16384)
    except (ImportError, OSError, pkg.PetstripError) as e:
        return {"error": str(e)}
    sec = kernel_ms * 1e-3
    n_lds = float(st["n_ellipse_pixels"])
    n_glob = float(st["n_window_pixels"])
    return {
        "lds_kind": lds_kind, "lds_per_iteration": n_lds, "lds_rate": n_lds / sec,
        "lds_peak": l_rate, "lds_frac": n_lds / sec / l_rate,
        "flush_kind": flush_kind,
        "global_per_iteration_max": n_glob, "global_rate_max": n_glob / sec,
        "global_peak": g_rate, "global_frac_max": n_glob / sec / g_rate,
        "updates_absorbed_in_lds": 1.0 - n_glob / max(n_lds, 1.0),
        "reference_global_per_iteration": n_lds,
        "reference_global_time_ms_at_peak": n_lds / g32 * 1e3,
    }


def hybrid_report(pkg, name, steps, no_cpu, cpu_threads, cpu_kind):
    """The 3D hybrid reconstruction (include/petstrip_hybrid.h) on a named
    workload: iteration time with the events resident on the device, parity
    of a 2e4-event prefix against the FP64 oracle after two iterations, and
    the reference's own iteration (oracle/_ref/libref_hybrid.so, OpenMP) on
    the same prefix as the CPU baseline."""
    import numpy as np
    g, lors, zzd, sz, sdl, length = pkg.hybrid_workload(name)
    t0 = time.perf_counter()
    with pkg.HybridReconstruction(g, sz, sdl, length) as rec:
        t_create = time.perf_counter() - t0
        t0 = time.perf_counter()
        rec.set_responses(lors, zzd)
        t_events = time.perf_counter() - t0
        st = rec.stats()
        rec(1)
        wall, kern = [], []
        for _ in range(steps):
            t0 = time.perf_counter()
            rec(1)
            wall.append((time.perf_counter() - t0) * 1e3)
            kern.append(rec.stats()["last_kernel_ms"])
    n_ev = st["n_events"]
    ms = float(np.median(wall))
    out = {
        "workload": name, "voxels": g.n_voxels, "pixel_infos": g.n_pixel_infos,
        "responses": len(lors), "events": n_ev,
        "voxels_per_event": st["n_voxels"] / max(n_ev, 1),
        "ms_per_iteration": ms, "kernel_ms": float(np.mean(kern)),
        "events_per_s_iter": n_ev / (ms * 1e-3),
        "voxel_evaluations_per_s": st["n_voxels"] / (float(np.mean(kern)) * 1e-3),
        "setup_s": {"weights_and_upload": t_create, "frame_events": t_events},
    }
    # the launch's bound: one float atomic per voxel (at most) against the
    # device's measured rate for the same kind of update over a buffer of the
    # accumulator's size (columns x rows x planes as a columns-wide image)
    try:
        peak, _ = pkg.probe_atomics(pkg.PROBE_GLOBAL_F32, g.n_columns,
                                    min(g.n_rows * g.n_planes, 32768), reps=256)
        achieved = st["n_voxels"] / (float(np.mean(kern)) * 1e-3)
        out["roofline"] = {"bound": "global float atomics", "achieved": achieved,
                           "peak": peak, "unit": "atomics/s", "frac": achieved / peak,
                           "kernel": "k_hybrid_mlem"}
    except (ImportError, OSError, pkg.PetstripError) as e:
        out["roofline"] = {"error": str(e)}
    if no_cpu:
        return out
    # the CPU leg: the oracle as the checker, the reference as the baseline
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests"))
    try:
        import oracle_lib as ol
    except ImportError as e:
        out["parity"] = {"error": str(e)}
        return out
    m = 20000
    with pkg.HybridReconstruction(g, sz, sdl, length) as rec:
        rec.set_responses(lors[:m], zzd[:m])
        w, sens = rec.weights()
        fr = rec.events()
        gpu = []
        for _ in range(2):
            rec(1)
            gpu.append(rec.rho)
    w64, s64 = ol.hybrid_weights(g, f64=True)
    o64 = ol.hybrid_mlem(g, sz, sdl, length, w64, s64, fr, 2, f64=True)
    out["parity"] = {
        "sample_events": len(fr),
        "rel_l2_vs_fp64": [float(np.linalg.norm(gpu[i] - o64[i]) / np.linalg.norm(o64[i]))
                           for i in range(2)]}
    if cpu_kind == "reference":
        mc = min(100000, len(lors))
        r = ol.ref_hybrid(g, sz, sdl, length, lors[:mc], zzd[:mc], 3, n_threads=cpu_threads,
                          keep_images=False)
        base = {"value": 3 * len(r["frames"]) / r["seconds"], "unit": "events/s/iter",
                "cores": cpu_threads, "kind": "reference", **host_info(),
                "sample": f"first {mc} responses, 3 iterations, mean iteration "
                          f"{r['seconds'] / 3:.3f} s, {cpu_threads} OpenMP threads"}
    else:
        w32, s32 = ol.hybrid_weights(g)
        fr32 = ol.hybrid_frames(g, sz, sdl, w32, lors[:m], zzd[:m])
        t0 = time.perf_counter()
        ol.hybrid_mlem(g, sz, sdl, length, w32, s32, fr32, 1)
        base = {"value": len(fr32) / (time.perf_counter() - t0), "unit": "events/s/iter",
                "cores": 1, "kind": "port", "sample": f"first {m} responses, one iteration"}
    out["cpu_baseline"] = base
    out["speedup_vs_cpu"] = out["events_per_s_iter"] / base["value"]
    return out


def time_gradient(rec, reps=3):
    """Median wall time (ms) of the full image + (sigma_z, sigma_dl) gradient
    of the Poisson list-mode log-likelihood at the current image: one launch
    computing lambda, the sigma derivative and the adjoint, kernels
    recomputed, plus the copy of the W*H gradient to the host."""
    import numpy as np
    import torch
    rec.backward()
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        g0 = time.perf_counter()
        rec.backward()
        torch.cuda.synchronize()
        times.append((time.perf_counter() - g0) * 1e3)
    return float(np.median(times))


def gradient_report(pkg, name):
    """The gradient against one MLEM iteration on the differentiation
    workload (256 x 256 pixels of 2 mm, 5e6 events of the three-ellipse
    phantom), events generated on the device, after two MLEM iterations."""
    import numpy as np
    n_pix, size, n_ev, _, phantom, seed = pkg.WORKLOADS[name]
    sc = pkg.workload_scanner(name)
    with pkg.Reconstruction(sc, device=0) as rec:
        rec.set_phantom_events(phantom, n_ev, seed=seed)
        rec(2)
        it_ms = []
        for _ in range(5):
            rec(1)
            it_ms.append(rec.stats()["last_iteration_ms"])
        g_ms = time_gradient(rec)
    mlem = float(np.median(it_ms))
    return {"workload": name, "events": n_ev, "mlem_iteration_ms": mlem,
            "grad_ms": g_ms, "grad_over_mlem": g_ms / mlem}


def variant_report(pkg, name, ev, no_cpu, threads):
    """The reference --gpu iteration as an option (--update gpu: the analytic
    kernel and the 5-point pixel sensitivity in the denominator,
    cuda/reconstruction.cu:59-113) on the same events: on the fused walks
    with the analytic kernel evaluated per pixel (the FP32 form), and for
    comparison on the per-event kernels (one wave per event over its box,
    FP64 image atomics).  Iteration times, events/s and, as the checker, one
    iteration on a prefix against the oracle's FP64 variant."""
    import numpy as np
    sc = pkg.workload_scanner(name)
    kw = dict(kernel=1, update=pkg.UPDATE_SENSITIVITY_DENOMINATOR)
    out = {"workload": name, "kernel": "analytic", "denominator": "sensitivity-weighted",
           "events": len(ev)}

    def timed(path):
        with pkg.Reconstruction(sc, device=0, path=path, **kw) as rec:
            rec.set_events(ev)
            fused = rec.stats()["fused"]
            rec(1)
            ms = []
            for _ in range(2):
                rec(1)
                ms.append(rec.stats()["last_iteration_ms"])
        return float(np.median(ms)), fused

    ms, fused = timed(pkg.PATH_AUTO)
    out["path"] = "fused walks, analytic kernel per pixel" if fused else "per-event kernels"
    out["ms_per_iteration"] = ms
    out["events_per_s_iter"] = len(ev) / (ms * 1e-3)
    out["per_event_ms_per_iteration"] = timed(pkg.PATH_PER_EVENT)[0]
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as ol
        m = min(20000, len(ev))
        sens = ol.sensitivity_f32(sc, 1).astype(np.float64)
        ref = ol.mlem_variant_f64(sc, ev[:m], 1, kernel=1, sens=sens, n_threads=threads)[0]
        with pkg.Reconstruction(sc, device=0, **kw) as rec:
            rec.set_events(ev[:m])
            rec(1)
            got = rec.rho.copy()
        l2, mx = _errors(got, ref)
        out["parity"] = {"rel_l2_vs_fp64": l2, "max_pixel_rel_vs_fp64": mx,
                         "sample": f"one iteration over the first {m} events"}
    return out


CFG5 = "strip-1024px-0.5mm-1e8ev"


def cfg5_report(pkg, rounds, threads, no_cpu):
    """The calibration workload (1024 x 1024 pixels of 0.5 mm, 1e8 events of
    the ten-ellipse phantom simulated at sigma = (0.01, 0.04), generated on
    the device): MLEM iteration and gradient times at the simulated sigma;
    `rounds` rounds of 5 MLEM iterations and a sigma step from (0.015, 0.06)
    with the support held at the start (calibration.py), its recovered
    sigma, events/s and, as the checker, one more iteration from the final
    image against FP64 on a prefix; and the same sigma ascent with the image
    held at the phantom's expected image of detected emissions, for at most
    six rounds, within which it settles (the estimator's check: it recovers
    the simulated widths)."""
    import numpy as np
    import torch
    n_pix, size, n_ev, _, phantom, seed = pkg.WORKLOADS[CFG5]
    sc = pkg.workload_scanner(CFG5)
    out = {"workload": CFG5, "events": n_ev}
    with pkg.Reconstruction(sc, device=0) as rec:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rec.set_phantom_events(phantom, n_ev, seed=seed)
        torch.cuda.synchronize()
        out["setup_s"] = time.perf_counter() - t0
        rec(1)
        it_ms = []
        for _ in range(3):
            rec(1)
            it_ms.append(rec.stats()["last_iteration_ms"])
        out["mlem_iteration_ms"] = float(np.median(it_ms))
        out["mlem_events_per_s_iter"] = n_ev / (out["mlem_iteration_ms"] * 1e-3)
        out["grad_ms"] = time_gradient(rec)
        out["grad_over_mlem"] = out["grad_ms"] / out["mlem_iteration_ms"]
        rec.rho = np.ones((n_pix, n_pix), np.float32)
        res = pkg.calibrate_sigma(
            rec, (0.015, 0.06), rounds=rounds,
            log=lambda h: print(f"cfg5 calibration round {h['round']}: sigma "
                                f"({h['sigma_z']:.6f}, {h['sigma_dl']:.6f}) loglik "
                                f"{h['loglik']:.6e}", file=sys.stderr, flush=True))
        state = (res["sigma_z"], res["sigma_dl"], rec.rho.copy())
        out["calibration"] = {
            "start": [0.015, 0.06], "simulated": [sc.sigma_z, sc.sigma_dl],
            "recovered": [res["sigma_z"], res["sigma_dl"]], "rounds": rounds,
            "mlem_iterations": res["mlem_iterations"], "support": "start",
            "seconds": res["seconds"],
            "mlem_events_per_s_per_iter": n_ev * res["mlem_iterations"] / res["seconds_mlem"],
            "gradient_s_per_round": res["seconds_gradient"] / max(rounds, 1),
            "trajectory": [[h["sigma_z"], h["sigma_dl"]] for h in res["history"]],
        }
        # the estimator on the true image: sigma steps only
        rec.set_sigma(sc.sigma_z, sc.sigma_dl)
        rec.rho = pkg.expected_image(sc, phantom, n_ev, rec.sensitivity(0))
        n_exp = max(1, min(6, rounds))
        exp = pkg.calibrate_sigma(rec, (0.015, 0.06), rounds=n_exp, mlem_per_round=0)
        out["calibration_expected_image"] = {
            "recovered": [exp["sigma_z"], exp["sigma_dl"]], "rounds": n_exp,
            "seconds": exp["seconds"]}
    if not no_cpu:
        ev = pkg.workload_events(CFG5, n=500_000)
        box_px = 23126 * (0.015 / 0.01) * (0.06 / 0.04)
        out["calibration"]["parity"] = calibration_parity(pkg, CFG5, ev, state, threads, box_px)
    return out


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="strip-512px-1mm-2e7ev")
    p.add_argument("--accumulation", default="fp32",
                   choices=["fp32", "compensated", "fixed64"])
    p.add_argument("--cpu-sample", type=float, default=None,
                   help="events in the CPU baseline sample (default: about "
                        "2.2e10 box pixels of work per iteration, 3.8e6 events at "
                        "512 px / 1 mm)")
    p.add_argument("--cpu-threads", type=int, default=None,
                   help="OpenMP threads of the CPU baseline (default: the CPUs in "
                        "this process's affinity mask, at most 16, the host share "
                        "of one GPU)")
    p.add_argument("--cpu-kind", default="reference", choices=["reference", "port"],
                   help="reference: the reference's own CPU path compiled from its "
                        "sources (oracle/_ref, built by __graft_entry__.build()); "
                        "port: the oracle's restatement")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--hybrid", default="hybrid-ring192-160px-4mm-128planes-1e6ev",
                   help="3D hybrid workload measured beside the headline line "
                        "('none' skips it)")
    p.add_argument("--no-grad", action="store_true")
    p.add_argument("--no-row-table", action="store_true",
                   help="form every box row in every iteration instead of reading "
                        "the row table (A/B of the table)")
    p.add_argument("--no-horner", action="store_true",
                   help="sum each row's denominator terms by advancing the kernel "
                        "chain instead of by Horner's rule (A/B of the pass-1 walk)")
    p.add_argument("--no-variant", action="store_true",
                   help="skip the reference --gpu variant (analytic kernel) line")
    p.add_argument("--grad-workload", default="strip-256px-2mm-5e6ev",
                   help="workload of the gradient-vs-MLEM-iteration line "
                        "('none' skips it)")
    p.add_argument("--calibrate", type=int, default=0, metavar="ROUNDS",
                   help="also run ROUNDS rounds of sigma calibration (5 MLEM "
                        "iterations + one gradient step on (sigma_z, sigma_dl) "
                        "each) from (0.015, 0.06) on the same events")
    p.add_argument("--cfg5-rounds", type=int, default=20,
                   help="rounds of the calibration block on the 1024 x 1024 / 0.5 mm "
                        "workload with 1e8 events (0 skips the block)")
    p.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                   help="also run one device copy of MIB MiB (for HBM counter "
                        "calibration under rocprofv3 --pmc)")
    return p.parse_args()


def measured_traffic(workload):
    """HBM bytes per MLEM iteration of the fused kernel (its launches: one
    per box height and window width class) from the committed rocprofv3
    --pmc summary (profiles/traffic.json), if it covers this workload."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    entry = t.get(workload)
    return None if entry is None else entry.get("bytes_per_iteration")


def _errors(img, ref):
    """Relative L2 error and largest per-pixel relative error (over pixels
    at least 1e-3 of the reference maximum) of img against ref."""
    import numpy as np
    a = np.asarray(img, np.float64).ravel()
    b = np.asarray(ref, np.float64).ravel()
    rel_l2 = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    m = b >= 1e-3 * b.max()
    return rel_l2, float(np.max(np.abs(a[m] - b[m]) / b[m]))


def host_info():
    """The CPUs this process may run on, nproc and the CPU model."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"affinity_cpus": len(os.sched_getaffinity(0)), "nproc": os.cpu_count(),
            "cpu_model": model}


def default_threads():
    """The CPUs in the affinity mask, at most 16: the host share of one GPU
    on the GPU boxes (their nproc counts the whole machine)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_baseline(sc, ev, n_sample, threads, kind, gpu_iterations, box_px):
    """The reference CPU path on a prefix of the events: iterations 1-3
    from rho = 1, each timed on its own, events/s/iter = events / the median
    of iterations 2-3; then, as the checker, the GPU's images on the same
    prefixes against the CPU image of iteration 1 and against FP64 runs of
    the same algorithm."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np

    import oracle_lib as ol
    sample = np.ascontiguousarray(ev[:int(n_sample)])
    fn = ol.ref_mlem if kind == "reference" else ol.mlem_f32
    fn(sc, sample[:min(len(sample), 20000)], 1, n_threads=threads)  # warm
    times, rho, cpu_img = [], None, None
    for it in range(3):
        t0 = time.perf_counter()
        rho = fn(sc, sample, 1, n_threads=threads, rho0=rho)[0]
        times.append(time.perf_counter() - t0)
        if it == 0:
            cpu_img = rho.copy()
    dt = float(np.median(times[1:]))
    base = {"value": len(sample) / dt, "unit": "events/s/iter", "cores": threads,
            "kind": kind, **host_info(),
            "iteration_s": times,
            "sample": f"first {len(sample)} events of the same event list, iterations "
                      f"1-3 from rho=1 each timed, median of iterations 2-3 "
                      f"({dt:.2f} s), {threads} OpenMP threads"}
    rel_cpu, max_cpu = _errors(gpu_iterations(sample, 1)[0], cpu_img)
    # FP64 runs: a prefix of about 3e9 box pixels, 2 iterations
    n64 = int(min(5e5, 3e9 / max(box_px, 1.0), len(sample)))
    it64 = ol.mlem_f64(sc, sample[:n64], 2, n_threads=threads)
    itg = gpu_iterations(sample[:n64], 2)
    errs = [_errors(g, r) for g, r in zip(itg, it64)]
    parity = {
        "rel_l2_vs_cpu": rel_cpu, "max_pixel_rel_vs_cpu": max_cpu,
        "cpu_sample": f"1 iteration over the first {len(sample)} events, against "
                      f"the {kind} CPU image",
        "rel_l2_vs_fp64": [e[0] for e in errs],
        "max_pixel_rel_vs_fp64": [e[1] for e in errs],
        "fp64_sample": f"iterations 1-2 over the first {n64} events; max-pixel error "
                       "over pixels >= 1e-3 of the FP64 maximum",
    }
    return base, parity


def calibration_parity(pkg, name, ev, state, threads, box_px):
    """The checker for the calibrated state: one iteration from the final
    image at the recovered sigmas, GPU against FP64, on a prefix of the
    events."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    sz, sdl, rho = state
    sc = pkg.workload_scanner(name, sigma_z=sz, sigma_dl=sdl)
    n = int(min(5e5, 1.5e9 / max(box_px, 1.0), len(ev)))
    sample = ev[:n]
    ref = ol.mlem_f64(sc, sample, 1, n_threads=threads, rho0=rho)[0]
    with pkg.Reconstruction(sc, device=0) as r:
        r.set_events(sample)
        r.rho = rho
        r(1)
        got = r.rho.copy()
    l2, mx = _errors(got, ref)
    return {"rel_l2_vs_fp64": l2, "max_pixel_rel_vs_fp64": mx,
            "sample": f"one iteration from the final image at the recovered sigmas "
                      f"over the first {n} events"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not "
              "supported (the path would shard events across GPUs and all-reduce "
              "the image)", file=sys.stderr)
        sys.exit(2)
    if args.cpu_threads is None:
        args.cpu_threads = default_threads()
    if not args.no_cpu and args.cpu_kind == "reference":
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as ol
        missing = [p for p, ok in ((ol.REF_SO, ol.have_ref()),
                                   (ol.REF_HYBRID_SO, ol.have_ref_hybrid())) if not ok]
        if missing:
            print("bench.py: the reference CPU baseline needs " + ", ".join(missing) +
                  " (built from the reference's sources by __graft_entry__.build() where "
                  "the reference tree exists); pass --cpu-kind port to time the oracle's "
                  "restatement instead", file=sys.stderr)
            sys.exit(2)
    import numpy as np
    import torch

    from __graft_entry__ import load_package
    pkg = load_package()
    name = args.workload
    n_pix, size, n_ev, _, _, _ = pkg.WORKLOADS[name]
    sc = pkg.workload_scanner(name)
    t0 = time.perf_counter()
    ev = pkg.workload_events(name)
    t_gen = time.perf_counter() - t0
    acc = {"fp32": pkg.ACCUM_FP32, "compensated": pkg.ACCUM_COMPENSATED,
           "fixed64": pkg.ACCUM_FIXED64}[args.accumulation]
    rec = pkg.Reconstruction(sc, device=0, accumulation=acc,
                             row_table=not args.no_row_table, horner=not args.no_horner)
    t0 = time.perf_counter()
    rec.set_events(ev)
    t_setup = time.perf_counter() - t0
    st = rec.stats()
    stream = torch.cuda.ExternalStream(rec.stream())
    rec(args.warmup)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    kernel_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        rec(1)
        ends[i].record(stream)
        kernel_ms.append(rec.stats()["last_kernel_ms"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    iter_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    value = n_ev * args.steps / elapsed
    k_ms = float(np.mean(kernel_ms))
    achieved = st["flop_alg"] / (k_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC, "value": value, "unit": "events/s/iter", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": name, "image": f"{n_pix}x{n_pix} @ {size * 1e3:g} mm",
                   "events": n_ev, "accumulation": args.accumulation,
                   "device": rec.device_name()},
        "roofline": {"bound": "valu", "achieved": achieved,
                     "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / VALU_PEAK_TFLOPS,
                     "traffic": measured_traffic(name),
                     "kernel": ("k_strip_fused<MLEM_K>" if args.accumulation == "fp32"
                                else "k_strip_fused<MLEM>"), "kernel_ms": k_ms,
                     "flop_per_launch": st["flop_alg"]},
        "iteration_ms": {"median": float(np.median(iter_ms)),
                         "min": float(np.min(iter_ms))},
        "row_table_bytes": st["row_table_bytes"],
        "work_items": {"total": st["n_work_items"],
                       "two_per_cu": st["n_items_two_per_cu"]},
        "footprint": {"box_px_per_event": st["n_box_pixels"] / max(n_ev, 1),
                      "ellipse_px_per_event": st["n_ellipse_pixels"] / max(n_ev, 1)},
        "setup_s": {"generate": t_gen, "upload_and_bin": t_setup},
    }
    out["atomics"] = atomic_report(pkg, st, n_pix, k_ms, args.accumulation)
    # the same events generated on the device (Philox stream of the host
    # generator) straight into a fresh handle's buffers, binned
    phantom, seed = pkg.WORKLOADS[name][4], pkg.WORKLOADS[name][5]
    with pkg.Reconstruction(sc, device=0, accumulation=acc) as rec_dev:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rec_dev.set_phantom_events(phantom, n_ev, seed=seed)
        torch.cuda.synchronize()
        out["setup_s"]["generate_on_device_and_bin"] = time.perf_counter() - t0
    if not args.no_grad:
        out["grad_ms"] = time_gradient(rec)
        out["grad_over_mlem"] = out["grad_ms"] / float(np.median(iter_ms))
    rec.close()
    if not args.no_grad and args.grad_workload != "none":
        out["gradient"] = gradient_report(pkg, args.grad_workload)
    cal_state = None
    if args.calibrate > 0:
        with pkg.Reconstruction(sc, device=0, accumulation=acc) as cal:
            cal.set_events(ev)
            res = pkg.calibrate_sigma(
                cal, (0.015, 0.06), rounds=args.calibrate,
                log=lambda h: print(f"calibration round {h['round']}: sigma "
                                    f"({h['sigma_z']:.6f}, {h['sigma_dl']:.6f}) "
                                    f"loglik {h['loglik']:.6e}", file=sys.stderr,
                                    flush=True))
            # the image before one more iteration at the recovered sigmas, for
            # the checker below
            cal_state = (res["sigma_z"], res["sigma_dl"], cal.rho.copy())
        out["calibration"] = {
            "start": [0.015, 0.06], "simulated": [sc.sigma_z, sc.sigma_dl],
            "recovered": [res["sigma_z"], res["sigma_dl"]],
            "rounds": args.calibrate, "mlem_iterations": res["mlem_iterations"],
            "seconds": res["seconds"],
            "mlem_events_per_s_per_iter": n_ev * res["mlem_iterations"] / res["seconds_mlem"],
            "gradient_s_per_round": res["seconds_gradient"] / args.calibrate,
            "trajectory": [[h["sigma_z"], h["sigma_dl"]] for h in res["history"]],
        }
    if args.calibrate_copy > 0:
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        out["calibration_copy_bytes"] = 2 * n * 4
        del a, b
    if not args.no_cpu:
        def gpu_iterations(events, n_iter):
            with pkg.Reconstruction(sc, device=0, accumulation=acc) as r:
                r.set_events(events)
                imgs = []
                for _ in range(n_iter):
                    r(1)
                    imgs.append(r.rho.copy())
            return imgs

        box_px = st["n_box_pixels"] / max(n_ev, 1)
        n_cpu = args.cpu_sample if args.cpu_sample else 2.2e10 / max(box_px, 1.0)
        out["cpu_baseline"], out["parity"] = cpu_baseline(
            sc, ev, n_cpu, args.cpu_threads, args.cpu_kind, gpu_iterations, box_px)
        if cal_state is not None:
            out["calibration"]["parity"] = calibration_parity(
                pkg, name, ev, cal_state, args.cpu_threads,
                st["n_box_pixels"] / max(n_ev, 1))
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    if args.cfg5_rounds > 0 and name != CFG5:
        print(f"calibration workload {CFG5}", file=sys.stderr, flush=True)
        out["cfg5"] = cfg5_report(pkg, args.cfg5_rounds, args.cpu_threads, args.no_cpu)
    if not args.no_variant:
        print(f"reference --gpu variant on {name}", file=sys.stderr, flush=True)
        out["variant"] = variant_report(pkg, name, ev, args.no_cpu, args.cpu_threads)
    if args.hybrid != "none":
        print(f"hybrid workload {args.hybrid}", file=sys.stderr, flush=True)
        out["hybrid"] = hybrid_report(pkg, args.hybrid, args.steps, args.no_cpu,
                                      args.cpu_threads, args.cpu_kind)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
