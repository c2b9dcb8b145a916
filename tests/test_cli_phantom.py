"""The 2d_strip_phantom command line (j-pet-mlem_amd/cli): description files,
options, the response / image / config outputs on host threads (-D -1), and
on the GPU the device run, the tan-bins maps and a reconstruction of its
output through the accompanying .cfg (src/2d/strip/phantom_cmd.cpp:136-444,
src/util/cmdline_hooks.cpp:110-123)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PH = os.path.join(ROOT, "j-pet-mlem_amd", "cli", "2d_strip_phantom")
REC = os.path.join(ROOT, "j-pet-mlem_amd", "cli", "2d_strip_reconstruction")
needs_cli = pytest.mark.skipif(not os.path.exists(PH), reason="CLI not built")

GEOM = ["-r", "0.43", "-l", "0.5", "-p", "0.004", "-n", "96", "--s-z", "0.01",
        "--s-dl", "0.04"]


def run(exe, args):
    return subprocess.run([exe] + args, capture_output=True, text=True, timeout=300)


def description(tmp_path):
    p = tmp_path / "phantom"
    p.write_text("ellipse -0.08 0.03 0.03 0.03 0 0\n"
                 "rectangle 0.06 -0.02 0.02 0.01 4\n"
                 "ellipse 0 0 0.15 0.1 30 1\n")
    return p


@needs_cli
def test_outputs_on_host_threads(tmp_path, pkg):
    d = description(tmp_path)
    out = tmp_path / "p.bin"
    r = run(PH, ["-D", "-1", "-v", "-e", "30k", "--seed", "7", "--full", "-o", str(out),
                 str(d)] + GEOM)
    assert r.returncode == 0, r.stderr
    assert " emitted: 30000 events" in r.stderr
    # the responses are the library generator's, emission mode
    sc = pkg.Scanner(0.43, 0.5, 96, 96, 0.004, 0.004, 0.01, 0.04)
    regs = [(0, -0.08, 0.03, 0.03, 0.03, 0, 0), (1, 0.06, -0.02, 0.02, 0.01, 0, 4),
            (0, 0, 0, 0.15, 0.1, 30, 1)]
    g = pkg.phantom_grid(96, 96, 0.004)
    want = pkg.phantom_generate(sc, regs, 30000, seed=7, grid=g, images=True, full=True)
    resp = np.fromfile(out, np.float32).reshape(-1, 3)
    assert np.array_equal(resp, want["responses"])
    assert f"detected: {len(resp)} events" in r.stderr
    assert np.array_equal(np.fromfile(tmp_path / "p_wo_error.bin", np.float32).reshape(-1, 3),
                          want["exact"])
    assert np.array_equal(np.fromfile(tmp_path / "p_events.bin", np.float32).reshape(-1, 4),
                          want["events"])
    for k in ("emitted", "detected", "naive"):
        img = np.fromfile(tmp_path / f"p_{k}", np.int32).reshape(96, 96)
        assert np.array_equal(img, want[k]), k
        hdr = (tmp_path / f"p_{k}.nrrd").read_text().splitlines()
        assert hdr == ["NRRD0001", "type: int", "endian: little", "dimension: 2",
                       "sizes: 96 96", f"data file: ./p_{k}", "encoding: raw"]
        assert (tmp_path / f"p_{k}.png").exists()
    cfg = (tmp_path / "p.cfg").read_text().splitlines()
    for line in ["--radius=0.43", "--length=0.5", "--s-pixel=0.004", "--n-z-pixels=96",
                 "--n-y-pixels=96", "--s-z=0.01", "--s-dl=0.04", "--scale=1"]:
        assert line in cfg, (line, cfg)


@needs_cli
def test_text_output_scale_point_and_errors(tmp_path, pkg):
    d = description(tmp_path)
    out = tmp_path / "t.txt"
    r = run(PH, ["-D", "-1", "-e", "5000", "--scale", "0.5", "--point", "0.01,0.02,0.5",
                 "--detected", "-o", str(out), str(d)] + GEOM)
    assert r.returncode == 0, r.stderr
    lines = out.read_text().splitlines()
    assert len(lines) == 5000 and len(lines[0].split()) == 3
    sc = pkg.Scanner(0.43, 0.5, 96, 96, 0.004, 0.004, 0.01, 0.04)
    regs = [(0, -0.04, 0.015, 0.015, 0.015, 0, 0), (1, 0.03, -0.01, 0.01, 0.005, 0, 4),
            (0, 0, 0, 0.075, 0.05, 30, 1), (2, 0.01, 0.02, 0, 0, 0, 0.5)]
    want = pkg.phantom_generate(sc, regs, 5000, seed=1, detected=True)["responses"]
    got = np.loadtxt(out, dtype=np.float64)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-7)
    bad = tmp_path / "bad"
    bad.write_text("hexagon 0 0 1\n")
    r = run(PH, ["-D", "-1", "-e", "10", str(bad)] + GEOM)
    assert r.returncode == 1 and "unknown phantom type: hexagon" in r.stderr
    r = run(PH, ["-D", "-1", "-e", "10", "--point", "1"] + GEOM)
    assert r.returncode == 1 and "--point" in r.stderr
    r = run(PH, ["-D", "-1", "-e", "10"] + GEOM)
    assert r.returncode == 1 and "phantom description expected" in r.stderr
    assert run(PH, []).returncode == 0


@needs_cli
@pytest.mark.gpu
def test_device_run_tan_maps_and_reconstruction(tmp_path, gpu_pkg):
    d = description(tmp_path)
    out = tmp_path / "g.bin"
    r = run(PH, ["-e", "200k", "--seed", "3", "--tan-bins", "8", "--kernel-point", "0.01,0",
                 "-o", str(out), str(d)] + GEOM)
    assert r.returncode == 0, r.stderr
    host = run(PH, ["-D", "-1", "-e", "200k", "--seed", "3", "-o", str(tmp_path / "h.bin"),
                    str(d)] + GEOM)
    assert host.returncode == 0, host.stderr
    a = np.fromfile(out, np.float32).reshape(-1, 3)
    b = np.fromfile(tmp_path / "h.bin", np.float32).reshape(-1, 3)
    assert a.shape == b.shape
    assert np.mean(np.all(a.view(np.uint32) == b.view(np.uint32), axis=1)) > 0.9999
    tb = np.fromfile(tmp_path / "g_tan_bins", np.int32).reshape(8, 96, 96)
    naive = np.fromfile(tmp_path / "g_naive", np.int32).reshape(96, 96)
    assert tb.sum() <= naive.sum() and tb.sum() > 0.9 * naive.sum()
    tk = np.fromfile(tmp_path / "g_tan_kernel", np.float32).reshape(8, 96, 96)
    assert np.all(np.isfinite(tk)) and tk.max() > 0
    hdr = (tmp_path / "g_tan_kernel.nrrd").read_text().splitlines()
    assert "type: float" in hdr and "sizes: 96 96 8" in hdr
    # the reconstruction reads g.cfg for the geometry
    rr = run(REC, [str(out), "-i", "1", "-I", "2", "-o", str(tmp_path / "rec.bin")])
    assert rr.returncode == 0, rr.stderr
    img = np.fromfile(tmp_path / "rec_2.bin", np.float32).reshape(96, 96)
    assert np.isfinite(img).all() and img.sum() == pytest.approx(len(a), rel=1e-3)
