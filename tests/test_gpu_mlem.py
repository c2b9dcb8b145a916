"""GPU MLEM parity against the CPU oracle (which is bitwise equal to the
reference CPU build, see test_oracle_reference.py).

Tolerances: discrete decisions (box, ellipse membership) must match exactly;
per-event denominators within 2e-6 relative (kernel value evaluated by the
device exp/sqrt instead of host libm); images within 1e-5 relative L2 per
iteration of the FP32 oracle and within 1e-4 of the FP64 oracle.
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) /
                 max(np.linalg.norm(np.asarray(b, np.float64)), 1e-300))


def small(gpu_pkg, n_pix=64, size=0.008, n=20000, seed=3, phantom=None):
    pkg = gpu_pkg
    sc = pkg.Scanner(0.43, 0.5, n_pix, n_pix, size, size, 0.01, 0.04)
    ev = pkg.phantom_events(sc, phantom or pkg.workloads.PHANTOM_3, n, seed=seed)
    return sc, ev


@pytest.mark.parametrize("row_scan", [False, True])
def test_footprints_exact(gpu_pkg, row_scan):
    """Per-event box and ellipse pixel counts of the row intervals the fused
    walks use (from the roots of each row's quadratic, or by scanning)
    against the oracle's test of every box pixel."""
    sc, ev = small(gpu_pkg, n=30000)
    with gpu_pkg.Reconstruction(sc, row_scan=row_scan) as rec:
        rec.set_events(ev)
        nb, ne = rec.footprints()
        st = rec.stats()
    onb, one = ol.footprint(sc, ev)
    print(f"case: 64x64, 3e4 events, row_scan {row_scan}, box {onb.sum()} "
          f"ellipse {one.sum()}", flush=True)
    assert np.array_equal(nb, onb)
    assert np.array_equal(ne, one)
    assert st["n_box_pixels"] == int(onb.sum())
    assert st["n_ellipse_pixels"] == int(one.sum())


def test_footprints_exact_512px_workload(gpu_pkg):
    """Row intervals at 512 x 512 pixels of 1 mm (2e5 events of the
    six-ellipse phantom, ~3900 ellipse pixels each) against the oracle."""
    name = "strip-512px-1mm-2e7ev"
    sc = gpu_pkg.workload_scanner(name)
    ev = gpu_pkg.workload_events(name, n=200000)
    with gpu_pkg.Reconstruction(sc) as rec:
        rec.set_events(ev)
        nb, ne = rec.footprints()
    onb, one = ol.footprint(sc, ev, n_threads=16)
    bad = np.flatnonzero(ne != one)
    print(f"case: {name} 2e5 events, ellipse pixels {one.sum()}, mismatches "
          f"{bad.size}", flush=True)
    assert np.array_equal(nb, onb)
    assert bad.size == 0


def test_row_scan_identical_image(gpu_pkg):
    """The deterministic MLEM image is bit-identical whether each row's
    interval comes from the root estimate or from the scan."""
    sc, ev = small(gpu_pkg, n=20000)
    outs = []
    for row_scan in (False, True):
        with gpu_pkg.Reconstruction(sc, accumulation=gpu_pkg.ACCUM_FIXED64,
                                    row_scan=row_scan) as rec:
            rec.set_events(ev)
            rec(2)
            outs.append(rec.rho)
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


def test_row_scan_identical_image_banded(gpu_pkg):
    """The same bit identity in the banded form, where the default options
    run kernels compiled for them (rows from their roots) and row_scan runs
    the general kernels: the deterministic image and the denominators."""
    sc, ev = small(gpu_pkg, n_pix=200, size=0.002, n=20000)
    outs = []
    for row_scan in (False, True):
        with gpu_pkg.Reconstruction(sc, accumulation=gpu_pkg.ACCUM_FIXED64,
                                    row_scan=row_scan, force_bands=1) as rec:
            rec.set_events(ev)
            assert rec.stats()["fused"] == 2
            rec(2)
            outs.append((rec.rho, rec.forward()))
    print(f"case: 200x200 @ 0.002 m, banded, {len(ev)} events", flush=True)
    assert np.array_equal(outs[0][0].view(np.uint32), outs[1][0].view(np.uint32))
    assert np.array_equal(outs[0][1].view(np.uint32), outs[1][1].view(np.uint32))


def test_denominators(gpu_pkg):
    sc, ev = small(gpu_pkg, n=20000)
    rng = np.random.default_rng(5)
    rho = rng.uniform(0.5, 2.0, (sc.n_y_pixels, sc.n_z_pixels)).astype(np.float32)
    with gpu_pkg.Reconstruction(sc) as rec:
        rec.set_events(ev)
        rec.rho = rho
        lam = rec.forward()
    ref = ol.denominators_f32(sc, ev, rho)
    pos = ref > 0
    assert np.array_equal(lam > 0, pos)
    err = np.abs(lam[pos] - ref[pos]) / ref[pos]
    print(f"denominator max rel err {err.max():.3e}", flush=True)
    assert err.max() < 2e-6


MODES = [("auto", "fp32"), ("auto", "compensated"), ("auto", "fixed64"),
         ("per_event", "fp32"), ("per_event", "compensated"), ("per_event", "fixed64")]


def opts(pkg, path, acc):
    return dict(path={"auto": pkg.PATH_AUTO, "per_event": pkg.PATH_PER_EVENT}[path],
                accumulation={"fp32": pkg.ACCUM_FP32,
                              "compensated": pkg.ACCUM_COMPENSATED,
                              "fixed64": pkg.ACCUM_FIXED64}[acc])


@pytest.mark.parametrize("path,acc", MODES)
@pytest.mark.parametrize("n_pix,size", [(64, 0.008), (41, 0.0125), (100, 0.005)])
def test_mlem_small(gpu_pkg, n_pix, size, path, acc):
    sc, ev = small(gpu_pkg, n_pix=n_pix, size=size)
    print(f"case: {n_pix}x{n_pix} @ {size} m, {len(ev)} events, {path}/{acc}", flush=True)
    with gpu_pkg.Reconstruction(sc, **opts(gpu_pkg, path, acc)) as rec:
        rec.set_events(ev)
        outs = []
        for _ in range(5):
            rec(1)
            outs.append(rec.rho)
    ref32 = ol.mlem_f32(sc, ev, 5, n_threads=8)
    ref64 = ol.mlem_f64(sc, ev, 5, n_threads=8)
    for i in range(5):
        e32, e64 = rel_l2(outs[i], ref32[i]), rel_l2(outs[i], ref64[i])
        print(f"iter {i + 1}: rel L2 vs f32 {e32:.2e} vs f64 {e64:.2e}", flush=True)
        assert e32 < 1e-5
        assert e64 < 1e-4
    # invariant: sum rho' = number of events with a positive denominator
    lam = ol.denominators_f32(sc, ev, ref32[3])
    assert abs(outs[4].sum(dtype=np.float64) - np.count_nonzero(lam > 0)) < 1e-3 * len(ev)
    # the last row and column are never visited (exclusive clamped box)
    assert np.all(outs[0][-1, :] == 0) and np.all(outs[0][:, -1] == 0)


def test_nonsquare_and_offsets(gpu_pkg):
    pkg = gpu_pkg
    sc = pkg.Scanner(0.45, 0.4, 70, 90, 0.006, 0.005, 0.012, 0.05, 0.01, -0.02)
    ev = pkg.phantom_events(sc, pkg.workloads.PHANTOM_3, 15000, seed=9)
    with pkg.Reconstruction(sc) as rec:
        rec.set_events(ev)
        rec(3)
        out = rec.rho
    osc = ol.SimpleScanner(0.45, 0.4, 70, 90, 0.006, 0.005, 0.012, 0.05, 0.01, -0.02)
    ref = ol.mlem_f32(osc, ev, 3, n_threads=8)[-1]
    assert out.shape == (70, 90)
    assert rel_l2(out, ref) < 1e-5


def test_edge_events(gpu_pkg):
    """Empty input, events outside the field of view and non-finite events
    contribute nothing and produce no NaN (reconstruction.h:290)."""
    pkg = gpu_pkg
    sc = pkg.Scanner(0.43, 0.5, 32, 32, 0.01, 0.01, 0.01, 0.04)
    with pkg.Reconstruction(sc) as rec:
        rec.set_events(np.zeros((0, 3), np.float32))
        rec(2)
        assert np.all(rec.rho == 0)
        bad = np.array([[5.0, 5.0, 0.0], [-7.0, -7.0, 0.1], [np.nan, 0.0, 0.0],
                        [np.inf, 0.0, 0.0], [0.0, 0.0, np.nan], [0.01, 0.02, 0.03]],
                       np.float32)
        rec.rho = np.ones((32, 32), np.float32)
        rec.set_events(bad)
        rec(1)
        out = rec.rho
        ref = ol.mlem_f32(ol.SimpleScanner(0.43, 0.5, 32, 32, 0.01, 0.01, 0.01, 0.04),
                          bad, 1)[0]
        assert np.all(np.isfinite(out))
        assert abs(out.sum() - 1.0) < 1e-5  # only the last event is inside
        assert rel_l2(out, ref) < 1e-6


def test_zero_denominator_count_after_new_events(gpu_pkg):
    """The count of events without a positive denominator (stats
    n_zero_denominator; it sets the next iteration's fixed-point unit) after
    each iterate call, over event sets of different sizes with and without
    events outside the strips on one handle: the denominators of the events
    outside the bins, which the binned iteration does not write, read 0."""
    pkg = gpu_pkg
    sc, ev = small(pkg, n=20000)
    far = ev.copy()
    far[::3, :2] = 5.0  # a third of the events far outside the strips
    with pkg.Reconstruction(sc) as rec:
        for evs in (ev, far, ev[:15000], far[:12000], ev):
            rec.set_events(evs)
            rec.rho = np.ones(rec.shape, np.float32)
            for k in range(2):
                rho_in = rec.rho
                rec(1)
                ref = ol.denominators_f32(sc, evs, rho_in)
                want = int(np.count_nonzero(~(ref > 0)))
                got = rec.stats()["n_zero_denominator"]
                print(f"case: {len(evs)} events, call {k}: {got} vs {want}", flush=True)
                assert got == want


def test_run_callbacks(gpu_pkg):
    sc, ev = small(gpu_pkg, n=5000)
    seen, prog = [], []
    with gpu_pkg.Reconstruction(sc) as rec:
        rec.set_events(ev)
        rho = np.ones((64, 64), np.float32)
        final = rec.run(rho, 2, 7, output=lambda it, img: seen.append((it, img.sum())),
                        progress=lambda c, f: prog.append((c, f)))
    its = [s[0] for s in seen]
    # sensitivity, first block 1-5, end of block 7, end of block 14
    assert its == [-1, 1, 2, 3, 4, 5, 7, 14]
    ref = ol.mlem_f32(sc, ev, 14, n_threads=8)
    assert rel_l2(final, ref[-1]) < 2e-5
    assert prog[-1] == (14, False)


def test_sensitivity(gpu_pkg):
    sc, _ = small(gpu_pkg, n=10)
    with gpu_pkg.Reconstruction(sc) as rec:
        for mode in (0, 1):
            s = rec.sensitivity(mode)
            ref = ol.sensitivity_f32(sc, mode)
            assert np.max(np.abs(s - ref)) < 2e-6


@pytest.mark.parametrize("path,variant", [("auto", False), ("per_event", False),
                                          ("per_event", True)])
def test_fixed64_bit_reproducible(gpu_pkg, path, variant):
    """The deterministic accumulation gives the same bits run after run: the
    fused walks and the per-event kernels (also with the reference --gpu
    variant's analytic kernel and sensitivity-weighted denominators)."""
    pkg = gpu_pkg
    sc, ev = small(pkg, n=30000)
    kw = opts(pkg, path, "fixed64")
    if variant:
        kw.update(kernel=1, update=pkg.UPDATE_SENSITIVITY_DENOMINATOR)
    outs = []
    for _ in range(2):
        with pkg.Reconstruction(sc, **kw) as rec:
            rec.set_events(ev)
            rec(3)
            outs.append(rec.rho)
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


@pytest.mark.parametrize("n_pix,size", [(64, 0.008), (100, 0.005), (200, 0.002)])
def test_row_table_identical(gpu_pkg, n_pix, size):
    """The row table (each binned event's box rows formed once per event set
    and kernel sigmas) gives the bits of the walks that form every row each
    iteration: the deterministic image, the denominators and, after a change
    of the kernel's sigmas with the support frozen (the table's kernel values
    rebuilt), the image again; the gradient and the FP32 image within their
    summation-order rounding.  Boxes of at most 16, 17-32
    and more than 32 rows (four, two and one events per wave)."""
    pkg = gpu_pkg
    sc, ev = small(pkg, n_pix=n_pix, size=size, n=30000)
    print(f"case: {n_pix}x{n_pix} @ {size} m, {len(ev)} events", flush=True)
    res = {}
    for table in (True, False):
        kw = opts(pkg, "auto", "fixed64")
        with pkg.Reconstruction(sc, row_table=table, **kw) as rec:
            rec.set_events(ev)
            st = rec.stats()
            assert st["fused"] == 1
            assert (st["row_table_bytes"] > 0) == table
            rec(3)
            out = [rec.rho, rec.forward()]
            rec.set_sigma(0.012, 0.045, freeze_support=True)
            rec.rho = np.ones(rec.shape, np.float32)
            rec(2)
            out.append(rec.rho)
            gr, gs, _ = rec.backward()
            out += [gr, np.asarray(gs)]
        with pkg.Reconstruction(sc, row_table=table) as rec:  # FP32
            rec.set_events(ev)
            rec(2)
            out.append(rec.rho)
        res[table] = out
    t, f = res[True], res[False]
    for i in range(3):
        assert np.array_equal(t[i].view(np.uint32), f[i].view(np.uint32)), i
    assert rel_l2(t[3], f[3]) < 1e-6
    # the sigma sums are FP64 sums over the events in their binned order,
    # which follows the scatter's atomics: equal up to their last bits
    np.testing.assert_allclose(t[4], f[4], rtol=1e-11)
    # the FP32 iteration rounds its values in a unit per work item, chosen
    # from the item's window, and the scatter's order decides which events
    # share an item: two runs differ by that rounding (~1e-6 here, ~3e5 adds
    # per pixel make it ~1e-7 at 512 x 512)
    assert rel_l2(t[5], f[5]) < 5e-6


@pytest.mark.parametrize("n_pix,size,bands", [(160, 0.004, False), (200, 0.002, False),
                                              (200, 0.002, True), (64, 0.008, False)])
def test_horner_walks_against_chain(gpu_pkg, n_pix, size, bands):
    """The iterations' pass 1 by Horner's rule against the chain walks
    (options pass1_walk = 1) on the same events: two FP32 iterations and the
    deterministic image; with bands also the denominators of the forward
    projection (its banded pass 1 is the iteration's); the gradient (whose
    walks keep the chain for the sigma moments) and the FP64 oracle's
    denominators unchanged."""
    pkg = gpu_pkg
    sc, ev = small(pkg, n_pix=n_pix, size=size, n=40000)
    rng = np.random.default_rng(5)
    rho = rng.uniform(0.2, 2.0, (n_pix, n_pix)).astype(np.float32)
    res = {}
    for horner in (True, False):
        with pkg.Reconstruction(sc, horner=horner, force_bands=bands) as rec:
            rec.set_events(ev)
            st = rec.stats()
            assert st["fused"] == (2 if bands else 1)
            assert st["horner_walks"] == int(horner)
            rec.rho = rho
            lam = rec.forward()
            g, gs, ll = rec.backward(want_loglik=True)
            rec.rho = np.ones(rec.shape, np.float32)
            rec(2)
            out = [lam, g, np.asarray(gs), ll, rec.rho]
        with pkg.Reconstruction(sc, horner=horner, force_bands=bands,
                                accumulation=pkg.ACCUM_FIXED64) as rec:
            rec.set_events(ev)
            rec(2)
            out.append(rec.rho)
        res[horner] = out
    print(f"case: {n_pix}x{n_pix} @ {size} m, bands {bands}", flush=True)
    (lt, gt, st_, llt, it, dt), (lf, gf, sf, llf, if_, df) = res[True], res[False]
    assert rel_l2(lt, lf) < 1e-6
    m = lf > 1e-3 * lf.max()
    assert np.max(np.abs(lt[m] - lf[m]) / lf[m]) < 2e-6
    assert rel_l2(gt, gf) < 1e-6
    np.testing.assert_allclose(st_, sf, rtol=1e-4)
    assert llt == pytest.approx(llf, rel=1e-8)
    # FP32 images of two runs: within the per-item rounding of the FP32
    # iteration (test_row_table_identical); the deterministic ones closer
    assert rel_l2(it, if_) < 5e-6
    assert rel_l2(dt, df) < 2e-6
    _, lam64, _, gs64 = ol.loglik_f64(sc, ev, rho, want_lambda=True)
    print(f"sigma gradient: {st_} (chain {sf}), oracle {gs64}", flush=True)
    assert rel_l2(lt, lam64) < 2e-6
    np.testing.assert_allclose(st_, gs64, rtol=2e-4)


def test_per_event_compensated_denominators(gpu_pkg):
    """The per-event kernels' compensated denominators (two-sum per lane and
    across the wave) are closer to the FP64 sums than the plain FP32 ones."""
    pkg = gpu_pkg
    sc, ev = small(pkg, n=20000)
    rho = np.random.default_rng(4).uniform(0.5, 2.0, (64, 64)).astype(np.float32)
    lams = {}
    for acc in ("fp32", "compensated"):
        with pkg.Reconstruction(sc, **opts(pkg, "per_event", acc)) as rec:
            rec.set_events(ev)
            rec.rho = rho
            lams[acc] = rec.forward()
    _, ref = ol.loglik_f64(sc, ev, rho, want_lambda=True, want_grad=False)[:2]
    pos = ref > 0
    err = {k: np.abs(v[pos] - ref[pos]) / ref[pos] for k, v in lams.items()}
    print(f"mean rel err vs FP64: fp32 {err['fp32'].mean():.2e}, compensated "
          f"{err['compensated'].mean():.2e}", flush=True)
    assert err["compensated"].mean() < err["fp32"].mean()
    assert err["compensated"].max() < 2e-6


def test_lambda_of_fused_iteration(gpu_pkg):
    """The fused iteration's denominators equal the forward projector's."""
    sc, ev = small(gpu_pkg, n=20000)
    with gpu_pkg.Reconstruction(sc) as rec:
        rec.set_events(ev)
        rec(2)
        lam = rec.forward()
    ref = ol.denominators_f32(sc, ev, ol.mlem_f32(sc, ev, 2, n_threads=8)[-1])
    pos = ref > 0
    err = np.abs(lam[pos] - ref[pos]) / ref[pos]
    assert err.max() < 2e-5


@pytest.mark.parametrize("acc", ["fp32", "fixed64"])
def test_workload_160px_1e6(gpu_pkg, acc):
    """Strip scanner R=0.43 m, L=0.5 m, 160x160 pixels of 4 mm, 1e6 events
    from the three-ellipse phantom, 10 iterations: every iteration within
    1e-5 relative L2 of the FP32 oracle and 1e-4 of the FP64 oracle."""
    pkg = gpu_pkg
    name = "strip-160px-4mm-1e6ev"
    sc = pkg.workload_scanner(name)
    ev = pkg.workload_events(name)
    print(f"case: {name} {acc}", flush=True)
    with pkg.Reconstruction(sc, **opts(pkg, "auto", acc)) as rec:
        rec.set_events(ev)
        outs = []
        for _ in range(10):
            rec(1)
            outs.append(rec.rho)
    ref32 = ol.mlem_f32(sc, ev, 10, n_threads=16)
    ref64 = ol.mlem_f64(sc, ev, 10, n_threads=16)
    for i in range(10):
        e32, e64 = rel_l2(outs[i], ref32[i]), rel_l2(outs[i], ref64[i])
        print(f"iter {i + 1}: rel L2 vs f32 {e32:.2e} vs f64 {e64:.2e}", flush=True)
        assert e32 < 1e-5
        assert e64 < 1e-4


@pytest.mark.parametrize("n_pix,size", [(64, 0.008), (256, 0.002)])
def test_banded_matches_unbanded(gpu_pkg, n_pix, size):
    """The banded forms (per-band partial denominators, combined in a fixed
    order), in row bands and in row x column bands, against the one-window
    form on the same events: MLEM images, denominators and the gradient.
    At 256 pixels the windows of the bins next to column 128 are cut by the
    column bands."""
    sc, ev = small(gpu_pkg, n_pix=n_pix, size=size, n=20000)
    rng = np.random.default_rng(11)
    rho = rng.uniform(0.5, 1.5, (sc.n_y_pixels, sc.n_z_pixels)).astype(np.float32)
    out = {}
    for fb in (0, 1, 2):
        with gpu_pkg.Reconstruction(sc, force_bands=fb) as rec:
            rec.set_events(ev)
            st = rec.stats()
            rec.rho = rho
            lam = rec.forward()
            grad, gsig, ll = rec.backward(want_loglik=True)
            rec.rho = rho
            rec(3)
            out[fb] = (st["fused"], lam, grad, gsig, ll, rec.rho)
    print(f"case: {n_pix}x{n_pix} @ {size * 1e3:g} mm, fused "
          f"{[out[f][0] for f in (0, 1, 2)]}", flush=True)
    assert [out[f][0] for f in (0, 1, 2)] == [1, 2, 3]
    a = out[0]
    for fb in (1, 2):
        b = out[fb]
        assert np.allclose(a[1], b[1], rtol=2e-6, atol=0)
        assert rel_l2(b[2], a[2]) < 1e-6
        assert b[3][0] == pytest.approx(a[3][0], rel=1e-5)
        assert b[3][1] == pytest.approx(a[3][1], rel=1e-5)
        assert b[4] == pytest.approx(a[4], rel=1e-7)
        # FP32 images: band items round in their own units (the FP32
        # iteration's per-item unit, test_row_table_identical)
        assert rel_l2(b[5], a[5]) < 5e-6


def test_banded_tall_boxes_vs_oracle(gpu_pkg):
    """0.5 mm pixels: the boxes are ~234 rows tall, the windows do not fit
    LDS and the fused passes run in bands; 3 MLEM iterations against the
    FP32 and FP64 oracles, and the log-likelihood gradient against FP64."""
    sc = gpu_pkg.Scanner(0.43, 0.5, 200, 200, 0.0005, 0.0005, 0.01, 0.04)
    ev = gpu_pkg.phantom_events(sc, gpu_pkg.workloads.PHANTOM_3, 6000, seed=21)
    with gpu_pkg.Reconstruction(sc) as rec:
        rec.set_events(ev)
        st = rec.stats()
        imgs = []
        for _ in range(3):
            rec(1)
            imgs.append(rec.rho)
        rho = imgs[-1]
        grad, gsig, _ = rec.backward()
    print(f"case: 200x200 @ 0.5 mm, 6000 events, fused {st['fused']}, "
          f"lds {st['lds_bytes']}", flush=True)
    assert st["fused"] == 2
    with gpu_pkg.Reconstruction(sc) as rec2:
        rec2.set_events(ev)
        nb, ne = rec2.footprints()
    onb, one = ol.footprint(sc, ev, n_threads=16)
    assert np.array_equal(nb, onb) and np.array_equal(ne, one)
    r32 = ol.mlem_f32(sc, ev, 3, n_threads=16)
    r64 = ol.mlem_f64(sc, ev, 3, n_threads=16)
    for i in range(3):
        e32, e64 = rel_l2(imgs[i], r32[i]), rel_l2(imgs[i], r64[i])
        print(f"iter {i + 1}: rel L2 vs f32 {e32:.2e} vs f64 {e64:.2e}", flush=True)
        assert e32 < 1e-5 and e64 < 1e-4
    _, _, g_ref, gs_ref = ol.loglik_f64(sc, ev, rho, n_threads=16)
    m = rho > 1e-3 * rho.max()
    assert rel_l2(grad[m], g_ref[m]) < 1e-5
    assert gsig[0] == pytest.approx(gs_ref[0], rel=1e-3)
    assert gsig[1] == pytest.approx(gs_ref[1], rel=1e-3)


def test_column_bands_wide_boxes_vs_oracle(gpu_pkg):
    """0.5 mm pixels at 1.5x the detector widths (sigma_z = 0.015 m,
    sigma_dl = 0.06 m, the start of a width calibration): boxes up to ~240
    columns wide, so even 64-row bands of the windows exceed LDS and the
    fused passes run in bands of 64 rows x 128 columns.  Footprints exact,
    2 MLEM iterations against the FP32 and FP64 oracles, the log-likelihood
    gradient against FP64."""
    pkg = gpu_pkg
    sc = pkg.Scanner(0.43, 0.5, 320, 320, 0.0005, 0.0005, 0.015, 0.06)
    phantom = [(0.0, 0.0, 0.05, 0.04, 20.0, 1.0), (0.02, 0.01, 0.015, 0.015, 0.0, 3.0)]
    ev = pkg.phantom_events(sc, phantom, 3000, seed=23)
    with pkg.Reconstruction(sc) as rec:
        rec.set_events(ev)
        st = rec.stats()
        nb, ne = rec.footprints()
        imgs = []
        for _ in range(2):
            rec(1)
            imgs.append(rec.rho)
        rho = imgs[-1]
        grad, gsig, _ = rec.backward()
    print(f"case: 320x320 @ 0.5 mm, sigma 1.5x, 3000 events, fused {st['fused']}, "
          f"lds {st['lds_bytes']}, items {st['n_work_items']}", flush=True)
    assert st["fused"] == 3
    onb, one = ol.footprint(sc, ev, n_threads=16)
    assert np.array_equal(nb, onb) and np.array_equal(ne, one)
    r32 = ol.mlem_f32(sc, ev, 2, n_threads=16)
    r64 = ol.mlem_f64(sc, ev, 2, n_threads=16)
    for i in range(2):
        e32, e64 = rel_l2(imgs[i], r32[i]), rel_l2(imgs[i], r64[i])
        print(f"iter {i + 1}: rel L2 vs f32 {e32:.2e} vs f64 {e64:.2e}", flush=True)
        assert e32 < 1e-5 and e64 < 1e-4
    _, _, g_ref, gs_ref = ol.loglik_f64(sc, ev, rho, n_threads=16)
    m = rho > 1e-3 * rho.max()
    assert rel_l2(grad[m], g_ref[m]) < 1e-5
    assert gsig[0] == pytest.approx(gs_ref[0], rel=1e-3)
    assert gsig[1] == pytest.approx(gs_ref[1], rel=1e-3)


def test_banded_pipelined_groups_vs_oracle(gpu_pkg):
    """600 x 600 pixels of 0.5 mm, 20000 events: over 4096 band items, so the
    banded passes run in several chunk groups, pass 2 of each on a second
    stream beside pass 1 of the next.  2 MLEM iterations against the FP32
    and FP64 oracles; the gradient against FP64."""
    pkg = gpu_pkg
    sc = pkg.Scanner(0.43, 0.5, 600, 600, 0.0005, 0.0005, 0.01, 0.04)
    ev = pkg.phantom_events(sc, pkg.workloads.PHANTOM_3, 20000, seed=31)
    with pkg.Reconstruction(sc) as rec:
        rec.set_events(ev)
        st = rec.stats()
        imgs = []
        for _ in range(2):
            rec(1)
            imgs.append(rec.rho)
        rho = imgs[-1]
        grad, gsig, _ = rec.backward()
        rec(1)
        nxt = rec.rho
    print(f"case: 600x600 @ 0.5 mm, 20000 events, fused {st['fused']}, "
          f"items {st['n_work_items']}", flush=True)
    assert st["fused"] == 2 and st["n_work_items"] >= 2 * 2048
    r32 = ol.mlem_f32(sc, ev, 2, n_threads=16)
    r64 = ol.mlem_f64(sc, ev, 2, n_threads=16)
    for i in range(2):
        e32, e64 = rel_l2(imgs[i], r32[i]), rel_l2(imgs[i], r64[i])
        print(f"iter {i + 1}: rel L2 vs f32 {e32:.2e} vs f64 {e64:.2e}", flush=True)
        assert e32 < 1e-5 and e64 < 1e-4
    # the MLEM identity (grad + 1) rho = rho' on the device's own iteration;
    # against FP64 the support can differ at the 3-sigma boundary (FP64
    # geometry), which at 0.5 mm shows at the 1e-5 level as for the images
    m = rho > 0
    assert rel_l2((grad + 1)[m] * rho[m], nxt[m]) < 1e-5
    _, _, g_ref, gs_ref = ol.loglik_f64(sc, ev, rho, n_threads=16)
    m = rho > 1e-3 * rho.max()
    assert rel_l2(grad[m], g_ref[m]) < 1e-4
    assert gsig[0] == pytest.approx(gs_ref[0], rel=1e-3)
    assert gsig[1] == pytest.approx(gs_ref[1], rel=1e-3)
