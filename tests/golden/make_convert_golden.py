"""Writes the 3d_tool_convert fixtures under tests/golden/convert/: seeded
synthetic inputs in the two formats the reference converter reads
(src/3d/tool/convert_cmd.cpp:88-210) and the reference converter's output
for each case, from oracle/_ref/3d_tool_convert (compiled from the
reference's own sources by oracle/Makefile).

Run in a container that holds the reference tree:
    make -C oracle && python tests/golden/make_convert_golden.py

Warsaw input lines hold x1 y1 z1 t1 x2 y2 z2 t2 (cm, ps), extra columns and
the scattering type last.  The hits are drawn on the detectors of the small
(24 x 9 mm x 15 mm, R 17.25 cm) and big (48 + 48 + 96 x 7 mm x 19 mm, two
rings rotated by half a detector) J-PET barrels, plus hits between
detectors, just outside a detector's face and past the scintillator's end.
T. Bednarski lines hold the two strip numbers (1-based) and the four times.
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "convert")
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "3d_tool_convert")

SMALL = dict(w=0.009, h=0.015, radius=[0.1725], n=[24], rotation=[0.0], length=0.3)
BIG = dict(w=0.007, h=0.019, radius=[0.4155, 0.458, 0.5655], n=[48, 48, 96],
           rotation=[0.0, 0.5, 0.5], length=0.5)


def detector_points(rng, geo, m):
    """m points (metres) on random detectors of geo, with a few deliberately
    between detectors or just outside a face."""
    rings = []
    for i, (r, n) in enumerate(zip(geo["radius"], geo["n"])):
        rot = geo["rotation"][i] if i > 0 else 0.0
        rings.append((r, n, rot))
    ring = rng.integers(0, len(rings), m)
    out = np.empty((m, 2))
    for k in range(m):
        r, n, rot = rings[ring[k]]
        d = rng.integers(0, n)
        ang = 2 * np.pi * d / n - np.pi / 2 + 2 * np.pi * rot / n
        kind = rng.random()
        if kind < 0.85:      # inside the detector
            u = rng.uniform(-0.49, 0.49) * geo["w"]
            v = rng.uniform(-0.49, 0.49) * geo["h"]
        elif kind < 0.93:    # just outside the inner face (within or past tol.)
            u = rng.uniform(-0.4, 0.4) * geo["w"]
            v = -geo["h"] / 2 - rng.choice([2e-5, 5e-5, 2e-4, 1e-3])
        else:                # between two detectors of the ring
            u = geo["w"] / 2 + rng.uniform(0.001, 0.004)
            v = rng.uniform(-0.4, 0.4) * geo["h"]
        x, y = u, r + geo["h"] / 2 + v
        c, s = np.cos(ang), np.sin(ang)
        out[k] = (x * c - y * s, x * s + y * c)
    return out


def warsaw_input(rng, geo, m, extra):
    p1 = detector_points(rng, geo, m) * 100.0
    p2 = detector_points(rng, geo, m) * 100.0
    half = geo["length"] * 50.0
    z1 = rng.uniform(-1.1, 1.1, m) * half
    z2 = rng.uniform(-1.1, 1.1, m) * half
    t1 = rng.uniform(500.0, 3000.0, m)
    t2 = rng.uniform(500.0, 3000.0, m)
    scatter = rng.integers(0, 4, m)
    lines = []
    for k in range(m):
        cols = [p1[k, 0], p1[k, 1], z1[k], t1[k], p2[k, 0], p2[k, 1], z2[k], t2[k]]
        cols += list(rng.uniform(0.0, 500.0, extra))
        lines.append(" ".join(f"{v:.6f}" for v in cols) + f" {scatter[k]}")
    return "\n".join(lines) + "\n"


def bednarski_input(rng, m):
    lines = []
    for _ in range(m):
        du, dd = rng.integers(1, 25, 2)
        t = rng.uniform(1000.0, 9000.0, 4)
        lines.append(f"{du} {dd} " + " ".join(f"{v:.3f}" for v in t))
    return "\n".join(lines) + "\n"


def geo_args(geo):
    return ["-w", str(geo["w"]), "-h", str(geo["h"]),
            "-r", ",".join(str(r) for r in geo["radius"]),
            "-d", ",".join(str(n) for n in geo["n"]),
            "--rotation", ",".join(str(r) for r in geo["rotation"]),
            "-l", str(geo["length"])]


# name: (input file, arguments)
CASES = {
    "warsaw_small": ("warsaw_small.txt", geo_args(SMALL) + ["-S", "7"]),
    "warsaw_small_only1": ("warsaw_small.txt",
                           geo_args(SMALL) + ["-S", "11", "--only", "1", "--warsaw"]),
    "warsaw_small_sigmas": ("warsaw_small.txt",
                            geo_args(SMALL) + ["--s-z", "0.01", "--s-dl", "0.04"]),
    "warsaw_big": ("warsaw_big.txt", geo_args(BIG) + ["-S", "3", "--warsaw"]),
    "warsaw_small_cfg": ("warsaw_small.txt", ["-c", "small_barrel.cfg", "-S", "5"]),
    "warsaw_auto_width": ("warsaw_small.txt", ["-r", "0.1725", "-d", "24", "-h", "0.015",
                                               "-l", "0.3", "-S", "2"]),
    "bednarski": ("bednarski.txt", ["--tbednarski", "-d", "24"]),
    "bednarski_relative": ("bednarski.txt", ["--tbednarski", "--relative-time", "-w", "0.009"]),
}

SMALL_CFG = """# small J-PET barrel: 24 detectors of 9 mm x 15 mm, inner radius 17.25 cm
--w-detector=0.009
--h-detector=0.015
--length=0.3
--radius=0.1725
--n-detectors=24
--fov-radius=0.150
# an option this command does not have: a warning only
--base-length=0.10
"""


def main():
    if not os.path.exists(REF):
        sys.exit(f"{REF} is missing: run make -C oracle in a container with the reference")
    os.makedirs(OUT, exist_ok=True)
    rng = np.random.default_rng(20261203)
    with open(os.path.join(OUT, "warsaw_small.txt"), "w") as f:
        f.write(warsaw_input(rng, SMALL, 3000, 4))
    with open(os.path.join(OUT, "warsaw_big.txt"), "w") as f:
        f.write(warsaw_input(rng, BIG, 3000, 2))
    with open(os.path.join(OUT, "bednarski.txt"), "w") as f:
        f.write(bednarski_input(rng, 500))
    with open(os.path.join(OUT, "small_barrel.cfg"), "w") as f:
        f.write(SMALL_CFG)
    for name, (inp, args) in CASES.items():
        r = subprocess.run([REF, *args, inp], cwd=OUT, capture_output=True, text=True,
                           timeout=120)
        if r.returncode != 0:
            sys.exit(f"{name}: reference failed: {r.stderr}")
        with open(os.path.join(OUT, name + ".out"), "w") as f:
            f.write(r.stdout)
        print(f"{name}: {len(r.stdout.splitlines())} lines", flush=True)


if __name__ == "__main__":
    main()
