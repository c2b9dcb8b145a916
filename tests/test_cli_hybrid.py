"""3d_hybrid_reconstruction (j-pet-mlem_amd/cli/hybrid_reconstruction.cpp):
the reference's --geometry command line (src/3d/hybrid/
reconstruction_cmd.cpp) over the hybrid C ABI.  The option checks run on the
CPU; the run itself on the GPU, against the Python host side on the same
geometry file and responses."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "j-pet-mlem_amd", "cli", "3d_hybrid_reconstruction")


def run(args, cwd):
    if not os.path.exists(TOOL):
        pytest.fail(f"{TOOL} is missing: build it with __graft_entry__.build()")
    return subprocess.run([TOOL, *args], cwd=cwd, capture_output=True, text=True, timeout=300)


def test_hybrid_cli_option_errors(tmp_path):
    r = run(["resp.txt"], tmp_path)
    assert r.returncode != 0 and "--geometry" in r.stderr
    r = run(["--system", "m.bin", "resp.txt"], tmp_path)
    assert r.returncode != 0 and "not supported" in r.stderr
    r = run(["--geometry", "missing.bin", "resp.txt"], tmp_path)
    assert r.returncode != 0 and "cannot open geometry" in r.stderr


def test_hybrid_cli_sensitivity_errors(pkg, tmp_path):
    """--sensitivity: a text volume (which the reference's reader leaves
    unset) and a short raw volume are refused before any device work."""
    hg = pkg.hybrid_geometry
    g = hg.ring_geometry(n_detectors=24, ring_radius=0.2, n_pixels=16, pixel_size=0.01)
    hg.write_geometry(g, str(tmp_path / "geom.bin"))
    (tmp_path / "resp.txt").write_text("")
    (tmp_path / "s.txt").write_text("1 1 1\n")
    r = run(["--geometry", "geom.bin", "--sensitivity", "s.txt", "resp.txt"], tmp_path)
    assert r.returncode != 0 and "text sensitivity" in r.stderr
    np.ones(16 * 16 * 8 - 1, np.float32).tofile(tmp_path / "s.bin")
    r = run(["--geometry", "geom.bin", "--sensitivity", "s.bin", "resp.txt"], tmp_path)
    assert r.returncode != 0 and "shorter than 8 planes" in r.stderr


@pytest.mark.gpu
def test_hybrid_cli_matches_library(gpu_pkg, tmp_path):
    pkg = gpu_pkg
    hg = pkg.hybrid_geometry
    g = hg.ring_geometry(n_detectors=48, ring_radius=0.22, n_pixels=32, pixel_size=0.0075)
    lors, zzd = hg.cylinder_responses(g, 6000, sigma_z=0.012, sigma_dl=0.05, length=0.3, seed=8)
    hg.write_geometry(g, str(tmp_path / "geom.bin"))
    with open(tmp_path / "resp.txt", "w") as f:
        for (a, b), (u, d, l) in zip(lors, zzd):
            f.write(f"{a} {b} {u:.6g} {d:.6g} {l:.6g}\n")
    # the responses as the reference's fast text reader reads them
    # (reconstruction.h:165-188; the reader is pinned bit for bit against
    # the reference in test_text_responses.py)
    zzd_txt, lors_txt = pkg.read_text_responses(tmp_path / "resp.txt", 5, with_lors=True)
    assert np.array_equal(lors_txt, lors)
    zzd = zzd_txt
    print("case: 32x32x32 ring geometry file, 6000 text responses, 2 blocks of 2", flush=True)
    r = run(["--geometry", "geom.bin", "--s-z", "0.012", "--s-dl", "0.05", "-l", "0.3",
             "-i", "2", "-I", "2", "-o", "out.bin", "resp.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    names = sorted(os.listdir(tmp_path))
    for n in ["out.bin", "out.nrrd", "out_1.bin", "out_2.bin", "out_4.bin", "out_4.nrrd"]:
        assert n in names, names
    assert "sizes: 32 32 32" in open(tmp_path / "out_4.nrrd").read()
    # the library on the same geometry file and responses (z_left from the
    # pixel size as the command line forms it)
    h = hg.read_geometry(str(tmp_path / "geom.bin"), n_planes=32, z_left=-32 * 0.0075 / 2)
    with pkg.HybridReconstruction(h, 0.012, 0.05, 0.3) as rec:
        rec.set_responses(lors, zzd)
        n_ev = rec.stats()["n_events"]
        rec(4)
        want = rec.rho
    got = np.fromfile(tmp_path / "out_4.bin", np.float32).reshape(want.shape)
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"  rel L2 command line vs library: {rel:.2e}", flush=True)
    # two runs differ only in the order of their float atomic sums
    assert rel < 1e-5
    naive = np.fromfile(tmp_path / "out.bin", np.int32)
    assert 0 < naive.sum() <= n_ev


@pytest.mark.gpu
def test_hybrid_cli_sensitivity_matches_library(gpu_pkg, tmp_path):
    """--sensitivity with a raw volume of n_planes / 2 planes against the
    library's set_sensitivity_3d on the same inputs."""
    pkg = gpu_pkg
    hg = pkg.hybrid_geometry
    g = hg.ring_geometry(n_detectors=48, ring_radius=0.22, n_pixels=32, pixel_size=0.0075)
    lors, zzd = hg.cylinder_responses(g, 6000, sigma_z=0.012, sigma_dl=0.05, length=0.3, seed=9)
    hg.write_geometry(g, str(tmp_path / "geom.bin"))
    with open(tmp_path / "resp.txt", "w") as f:
        for (a, b), (u, d, l) in zip(lors, zzd):
            f.write(f"{a} {b} {u:.9g} {d:.9g} {l:.9g}\n")
    vol = np.random.default_rng(9).uniform(0.5, 1.5, (16, 32, 32)).astype(np.float32)
    vol.tofile(tmp_path / "sens.bin")
    print("case: 32x32x32, 16-plane sensitivity volume, 6000 responses, 2 iterations",
          flush=True)
    r = run(["--geometry", "geom.bin", "--s-z", "0.012", "--s-dl", "0.05", "-l", "0.3",
             "--sensitivity", "sens.bin", "-i", "1", "-I", "2", "-o", "out.bin", "resp.txt"],
            tmp_path)
    assert r.returncode == 0, r.stderr
    h = hg.read_geometry(str(tmp_path / "geom.bin"), n_planes=32, z_left=-32 * 0.0075 / 2)
    with pkg.HybridReconstruction(h, 0.012, 0.05, 0.3) as rec:
        rec.set_sensitivity_3d(vol)
        rec.set_responses(lors, zzd)
        rec(2)
        want = rec.rho
    got = np.fromfile(tmp_path / "out_2.bin", np.float32).reshape(want.shape)
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"  rel L2 command line vs library: {rel:.2e}", flush=True)
    assert rel < 1e-5
