"""The 2d_strip_reconstruction command line (j-pet-mlem_amd/cli): argument
handling on the CPU, end-to-end runs and output files on the GPU.

File conventions checked against the reference's writers:
src/util/png_writer.h:31-54 (inverted 8-bit PNG), nrrd_writer.h:18-27,
cmdline_types.h:35-55 (index suffix), reconstruction_cmd.cpp:232-291.
"""
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "j-pet-mlem_amd", "cli", "2d_strip_reconstruction")

needs_cli = pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")


def run(args, **kw):
    return subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120, **kw)


def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w = 8, b"", None
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert zlib.crc32(typ + body) & 0xffffffff == crc
        if typ == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert depth == 8 and ctype == 0
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, w + 1)
    assert np.all(rows[:, 0] == 0)
    return rows[:, 1:]


@needs_cli
def test_usage_and_errors():
    r = run([])
    assert r.returncode == 0 and "usage" in r.stderr
    r = run(["/nonexistent/events.bin", "-i", "1"])
    assert r.returncode == 1 and "cannot open" in r.stderr
    r = run(["--bogus-option", "x.bin"])
    assert r.returncode == 1 and "undefined option" in r.stderr
    r = run(["-n", "12x", "x.bin"])
    assert r.returncode == 1


@needs_cli
def test_no_device_is_an_error_not_a_crash(tmp_path, pkg):
    if pkg.device_count() > 0:
        pytest.skip("a device is visible")
    ev = tmp_path / "ev.bin"
    np.zeros((4, 3), np.float32).tofile(ev)
    r = run([str(ev), "-i", "1", "-n", "16", "-p", "0.01"])
    assert r.returncode == 1 and "device" in r.stderr


def make_input(tmp_path, pkg, n=20000, txt=False):
    sc = pkg.Scanner(0.43, 0.5, 64, 64, 0.008, 0.008, 0.01, 0.04)
    ev = pkg.phantom_events(sc, pkg.workloads.PHANTOM_3, n, seed=3)
    if txt:
        path = tmp_path / "p_strip.txt"
        np.savetxt(path, ev, fmt="%.9g")
    else:
        path = tmp_path / "p_strip.bin"
        ev.tofile(path)
    # accompanying config, as 2d_strip_phantom writes it
    (tmp_path / "p_strip.cfg").write_text(
        "--radius=0.43\n--length=0.5\n--s-pixel=0.008\n--n-z-pixels=64\n"
        "--n-y-pixels=64\n# comment\n--s-z=0.01\n--s-dl=0.04\n")
    return sc, ev, path


@needs_cli
@pytest.mark.gpu
def test_gpu_branch_outputs(tmp_path, gpu_pkg):
    sc, ev, path = make_input(tmp_path, gpu_pkg)
    out = tmp_path / "r.bin"
    r = run([str(path), "-G", "-i", "2", "-I", "3", "-o", str(out), "-v", "-n", "64",
             "-p", "0.008", "-r", "0.43", "-l", "0.5", "--s-z", "0.01", "--s-dl", "0.04"])
    assert r.returncode == 0, r.stderr
    assert "# read 20000 responsess" in r.stderr and "# device:" in r.stderr
    names = sorted(os.listdir(tmp_path))
    for stem in ["r_sensitivity", "r_1", "r_2", "r_3", "r_6"]:
        for e in [".png", ".bin", ".nrrd"]:
            assert stem + e in names, (stem, e, names)
    assert "r_4.bin" not in names  # first block stops outputs before its end
    img = np.fromfile(tmp_path / "r_6.bin", np.float32).reshape(64, 64)
    ref = ol.mlem_f32(sc, ev, 6, n_threads=8)[-1]
    assert np.linalg.norm(img - ref) / np.linalg.norm(ref) < 1e-5
    nrrd = (tmp_path / "r_6.nrrd").read_text().splitlines()
    assert nrrd == ["NRRD0001", "type: float", "endian: little", "dimension: 2",
                    "sizes: 64 64", "data file: ./r_6.bin", "encoding: raw"]
    png = read_png(tmp_path / "r_6.png")
    want = np.maximum(0.0, 255.0 - (255.0 / img.max()) * img.astype(np.float64))
    assert np.max(np.abs(png.astype(int) - want.astype(np.uint8).astype(int))) <= 1


@needs_cli
@pytest.mark.gpu
def test_cpu_branch_outputs_and_config(tmp_path, gpu_pkg):
    sc, ev, path = make_input(tmp_path, gpu_pkg, txt=True)
    out = tmp_path / "o.txt"
    # geometry comes from the accompanying p_strip.cfg
    r = run([str(path), "-i", "2", "-I", "5", "-o", str(out)])
    assert r.returncode == 0, r.stderr
    names = sorted(os.listdir(tmp_path))
    assert "o_sensitivity.png" in names and "o_05.txt" in names and "o_10.txt" in names
    tup = np.loadtxt(tmp_path / "o_10.txt")
    img = np.zeros((64, 64))
    img[tup[:, 1].astype(int), tup[:, 0].astype(int)] = tup[:, 2]
    # the events as the reference's text reader reads them (pinned bit for
    # bit against the reference in test_text_responses.py)
    ev32 = gpu_pkg.read_text_responses(path)
    ref = ol.mlem_f32(sc, ev32, 10, n_threads=8)[-1]
    assert np.linalg.norm(img - ref) / np.linalg.norm(ref) < 1e-5


@needs_cli
@pytest.mark.gpu
def test_rho_resume(tmp_path, gpu_pkg):
    sc, ev, path = make_input(tmp_path, gpu_pkg)
    r = run([str(path), "-G", "-i", "1", "-I", "2", "-o", str(tmp_path / "a.bin")])
    assert r.returncode == 0, r.stderr
    r = run([str(path), "-G", "-i", "1", "-I", "1", "--rho", str(tmp_path / "a_2.bin"),
             "-o", str(tmp_path / "b.bin")])
    assert r.returncode == 0, r.stderr
    img = np.fromfile(tmp_path / "b_1.bin", np.float32).reshape(64, 64)
    ref = ol.mlem_f32(sc, ev, 3, n_threads=8)[-1]
    assert np.linalg.norm(img - ref) / np.linalg.norm(ref) < 1e-5


@needs_cli
@pytest.mark.gpu
def test_update_gpu_is_reference_gpu_iteration(tmp_path, gpu_pkg):
    """--update gpu: analytic kernel with the 5-point sensitivity in the
    denominator, against the oracle's FP64 statement of the reference --gpu
    iteration."""
    sc, ev, path = make_input(tmp_path, gpu_pkg)
    out = tmp_path / "g.bin"
    r = run([str(path), "-G", "-i", "1", "-I", "3", "-o", str(out), "--update", "gpu",
             "-n", "64", "-p", "0.008", "-r", "0.43", "-l", "0.5", "--s-z", "0.01",
             "--s-dl", "0.04"])
    assert r.returncode == 0, r.stderr
    img = np.fromfile(tmp_path / "g_3.bin", np.float32).reshape(64, 64)
    sens = ol.sensitivity_f32(sc, 1).astype(np.float64)
    ref = ol.mlem_variant_f64(sc, ev, 3, kernel=1, sens=sens, n_threads=8)[-1]
    assert np.linalg.norm(img - ref) / np.linalg.norm(ref) < 1e-5
    bad = run([str(path), "--update", "sideways", "-n", "64"])
    assert bad.returncode != 0


GOLD_TEXT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                         "text_responses.npz")
SCANNER_ARGS = ["-n", "64", "-p", "0.008", "-r", "0.43", "-l", "0.5", "--s-z", "0.01",
                "--s-dl", "0.04"]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@needs_cli
@pytest.mark.gpu
def test_text_input_reference_floats(tmp_path, gpu_pkg):
    """A .txt input of 6000 events printed with 6 significant digits (as
    the converter prints) is reconstructed from the floats the reference's
    fast text reader gives (tests/golden/text_responses.npz, from the
    reference compiled from its headers), not from strtod's; and "<v/100>e2"
    lines read as v/100, the reader's exponent rule (text_parser.h:83-91,
    168-169), which strtod would read as v."""
    gold = np.load(GOLD_TEXT)
    sc = gpu_pkg.Scanner(0.43, 0.5, 64, 64, 0.008, 0.008, 0.01, 0.04)
    for name in ["phantom", "exp2"]:
        path = tmp_path / f"{name}.txt"
        path.write_bytes(gold[f"file_{name}"].tobytes())
        print(f"case: 2d_strip_reconstruction -G on {name}.txt, 64x64, 3 iterations",
              flush=True)
        r = run([str(path), "-G", "-i", "1", "-I", "3", "-o", str(tmp_path / f"{name}.bin"),
                 "-v", *SCANNER_ARGS])
        assert r.returncode == 0, r.stderr
        n = int(gold[f"n_{name}"])
        assert f"# read {n} responsess" in r.stderr
        img = np.fromfile(tmp_path / f"{name}_3.bin", np.float32).reshape(64, 64)
        ref = ol.mlem_f32(sc, gold[f"events_{name}"], 3, n_threads=8)[-1]
        rel = _rel(img, ref)
        print(f"  rel L2 vs oracle on the reference reader's floats: {rel:.2e}", flush=True)
        assert rel < 1e-5
        if name == "exp2":
            text = gold["file_exp2"].tobytes().decode().split()
            strtod = np.array([float(t) for t in text], np.float32).reshape(-1, 3)
            wrong = ol.mlem_f32(sc, strtod, 3, n_threads=8)[-1]
            assert _rel(img, wrong) > 0.5


@needs_cli
@pytest.mark.gpu
def test_3d_response_scale_length_dl_is_time(tmp_path, gpu_pkg):
    """--3d-response on the reference converter's output (5 fields, the LOR
    fields read and dropped), then --scale-length and --dl-is-time as
    reconstruction_cmd.cpp:194-210 applies them: z_u, z_d (and dl unless it
    is a time) times the scale in double, rounded to float; a time dl times
    the speed of light times --scale-time.  The scaling runs after every
    input file over all responses read so far, so with two files the first
    file's events are scaled twice, as there."""
    gold = np.load(GOLD_TEXT)
    ev = gold["events_convert"].astype(np.float32)
    path = tmp_path / "conv.txt"
    path.write_bytes(gold["file_convert"].tobytes())
    sc = gpu_pkg.Scanner(0.43, 0.5, 64, 64, 0.008, 0.008, 0.01, 0.04)

    def scaled(e, s, time=None):
        e = e.copy()
        e[:, 0] = (e[:, 0].astype(np.float64) * s).astype(np.float32)
        e[:, 1] = (e[:, 1].astype(np.float64) * s).astype(np.float32)
        if time is None:
            e[:, 2] = (e[:, 2].astype(np.float64) * s).astype(np.float32)
        else:
            c, st = time
            e[:, 2] = ((e[:, 2].astype(np.float64) * c) * st).astype(np.float32)
        return e

    cases = [
        ("plain", [], ev),
        ("scale", ["--scale-length", "0.9"], scaled(ev, 0.9)),
        ("time", ["--scale-length", "0.9", "--dl-is-time", "--scale-time", "2.5e-9",
                  "--speed-of-light", "299792458"],
         scaled(ev, 0.9, (299792458.0, 2.5e-9))),
    ]
    for name, extra, want_ev in cases:
        print(f"case: --3d-response {name} {' '.join(extra)}", flush=True)
        if name == "time":
            # dl was written in metres: read it back as a time in units of
            # 1 / (c * scale-time) metres
            t = gold["file_convert"].tobytes().decode().splitlines()
            rows = [ln.split() for ln in t]
            dl_t = [f"{float(r[4]) / (299792458.0 * 2.5e-9):.6g}" for r in rows]
            text = "\n".join(" ".join(r[:4] + [d]) for r, d in zip(rows, dl_t)) + "\n"
            tpath = tmp_path / "conv_time.txt"
            tpath.write_text(text)
            base = gpu_pkg.read_text_responses(tpath, 5)
            want_ev = scaled(base, 0.9, (299792458.0, 2.5e-9))
            inp = tpath
        else:
            inp = path
        r = run([str(inp), "--3d-response", "-G", "-i", "1", "-I", "2",
                 "-o", str(tmp_path / f"{name}.bin"), *extra, *SCANNER_ARGS])
        assert r.returncode == 0, r.stderr
        img = np.fromfile(tmp_path / f"{name}_2.bin", np.float32).reshape(64, 64)
        ref = ol.mlem_f32(sc, want_ev, 2, n_threads=8)[-1]
        rel = _rel(img, ref)
        print(f"  rel L2 vs oracle: {rel:.2e}", flush=True)
        assert rel < 1e-5
    # two input files: the first file's events are scaled twice
    r = run([str(path), str(path), "--3d-response", "-G", "-i", "1", "-I", "1",
             "-o", str(tmp_path / "two.bin"), "--scale-length", "0.9", *SCANNER_ARGS])
    assert r.returncode == 0, r.stderr
    want_ev = np.concatenate([scaled(scaled(ev, 0.9), 0.9), scaled(ev, 0.9)])
    img = np.fromfile(tmp_path / "two_1.bin", np.float32).reshape(64, 64)
    ref = ol.mlem_f32(sc, want_ev, 1, n_threads=8)[-1]
    assert _rel(img, ref) < 1e-5
    # a .bin input with --3d-response is refused, as in the reference
    ev.tofile(tmp_path / "conv.bin")
    r = run([str(tmp_path / "conv.bin"), "--3d-response", "-i", "1", *SCANNER_ARGS])
    assert r.returncode != 0 and ".txt" in r.stderr


@needs_cli
def test_text_input_bad_line(tmp_path):
    """A line the reader cannot parse stops the command with the line in the
    message, before any device work (reconstruction.h:188-191)."""
    p = tmp_path / "bad.txt"
    p.write_text("0.1 0.2 0.3\n0.1\t0.2 0.3\n")
    r = run([str(p), "-i", "1", *SCANNER_ARGS])
    assert r.returncode == 1 and "error line: 0.1\t0.2 0.3" in r.stderr
