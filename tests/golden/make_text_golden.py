"""Writes tests/golden/text_responses.npz: what the reference's fast text
reader (util::text_parser, src/util/text_parser.h:29-175, through
Strip::Reconstruction::fast_load_txt_events, src/2d/strip/
reconstruction.h:169-196) reads from a set of number strings and response
files, taken from the reference itself, compiled from its own headers into
oracle/_ref/libref_strip.so (oracle/ref_driver.cpp: ref_text_parse_fields,
ref_fast_load_txt_events).

Run in a container that holds the reference tree:
    make -C oracle && python tests/golden/make_text_golden.py

Contents (all data; every input is synthetic and seeded here):
  num_strings / num_bits / num_ok  number strings (0-11 fraction digits,
      signs, exponents of both signs, whole parts and fractions past 2^31,
      malformed ones) and the float bits the reader gives (ok = 0: it
      throws); exponents of magnitude 300 are left out, where the reader
      indexes one past its 300-entry power tables
  int_strings / int_values         integer fields as operator>><int> reads
      them (the LOR indices of the 5-field form)
  file_<name> / events_<name> / n_<name> / cols_<name>
      response files as bytes and the float32 (z_u, z_d, dl) triples the
      reader returns for them, n = -1 when it throws:
        quirks     spaces, extra fields, CR before the newline, a last line
                   without a newline (not read)
        badline    a line that does not parse
        phantom    6000 phantom events printed with 6 significant digits,
                   as the converter prints (src/3d/tool/convert_cmd.cpp:
                   158-160), and some with 9 and in exponent form
        exp2       1000 phantom events printed as "<v / 100>e2": the reader
                   multiplies by 10 * 0.1f for e = 2 (text_parser.h:83-91),
                   so it reads v / 100
        convert    the reference converter's --3d-response output
                   (tests/golden/convert/warsaw_big.out), 5 fields
"""
import ctypes as C
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "libref_strip.so")


def load_ref():
    lib = C.CDLL(REF)
    lib.ref_text_parse_fields.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_void_p,
                                          C.c_void_p]
    lib.ref_text_parse_fields.restype = C.c_int
    lib.ref_fast_load_txt_events.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_int64]
    lib.ref_fast_load_txt_events.restype = C.c_int64
    return lib


def number_strings(rng):
    out = []
    for nd in range(0, 12):
        for _ in range(120):
            mag = 10.0 ** rng.integers(-3, 3)
            out.append(f"{rng.uniform(-1, 1) * mag:.{nd}f}")
    for _ in range(200):
        v = rng.uniform(-1, 1) * 10.0 ** rng.integers(-8, 8)
        out.append(f"{v:.6g}")
        out.append(f"{v:.9g}")
        out.append(f"{v:e}")
        out.append(f"{v:.3E}")
    for e in list(range(1, 40)) + [45, 46, 100, 299]:
        out += [f"1.5e{e}", f"2.5e-{e}", f"-3e+{e}", f"0.0025E{e}"]
    out += ["0", "-0", "+0", "0.0", "-0.0", ".5", "-.5", "+.25", "5.", "-5.",
            "00012.50000", "1234567.891", "16777217", "16777217.5", "2147483647",
            "2147483648", "4294967296.25", "99999999999.5", "0.12345678901",
            "0.9999999999", "3.14159265358979", "1e0", "1e-0", "1e+0", "7e-45",
            "1e-46", "2e-299", "12e5", "-0.000001", "123456789012.0", "1.5 2.5",
            "1.5abc", "0.25\r", "1e", "1e+", "e5", ".", "-", "+", "", "abc", "-.e1",
            "1.5e301", "0x10", "inf", "nan", " 7.25"]
    return out


def int_strings():
    return ["0", "12", "-7", "+7", "191", "32767", "32768", "70000", "-40000",
            "2147483647", "2147483648", "4294967301", "", "x", "-", "12.5"]


def parse_numbers(ref, strs):
    bits = np.zeros(len(strs), np.uint32)
    ok = np.zeros(len(strs), np.uint8)
    kinds = np.ones(1, np.int32)
    for k, s in enumerate(strs):
        f = np.zeros(1, np.float32)
        i = np.zeros(1, np.int32)
        rc = ref.ref_text_parse_fields(s.encode(), 1, kinds.ctypes.data, f.ctypes.data,
                                       i.ctypes.data)
        ok[k] = rc == 0
        bits[k] = f.view(np.uint32)[0] if rc == 0 else 0
    return bits, ok


def parse_ints(ref, strs):
    vals = np.zeros(len(strs), np.int32)
    kinds = np.zeros(1, np.int32)
    for k, s in enumerate(strs):
        f = np.zeros(1, np.float32)
        i = np.zeros(1, np.int32)
        assert ref.ref_text_parse_fields(s.encode(), 1, kinds.ctypes.data, f.ctypes.data,
                                         i.ctypes.data) == 0
        vals[k] = i[0]
    return vals


def load_file(ref, data, cols):
    with tempfile.NamedTemporaryFile(suffix=".txt", delete=False) as f:
        f.write(data)
        path = f.name
    try:
        cap = data.count(b"\n") + 1
        out = np.zeros((cap, 3), np.float32)
        n = ref.ref_fast_load_txt_events(path.encode(), int(cols == 5), out.ctypes.data, cap)
    finally:
        os.unlink(path)
    return (out[:n] if n >= 0 else np.zeros((0, 3), np.float32)), n


def phantom_events(n, seed):
    sys.path.insert(0, ROOT)
    from __graft_entry__ import load_package
    pkg = load_package()
    sc = pkg.Scanner(0.43, 0.5, 64, 64, 0.008, 0.008, 0.01, 0.04)
    return pkg.phantom_events(sc, pkg.workloads.PHANTOM_3, n, seed=seed)


def main():
    ref = load_ref()
    rng = np.random.default_rng(20261204)
    nums = number_strings(rng)
    bits, ok = parse_numbers(ref, nums)
    ints = int_strings()
    ivals = parse_ints(ref, ints)
    files = {}
    files["quirks"] = (b"0.1 0.2 0.3\n  -0.05   0.125 1e-2  extra fields\n"
                       b"0.25 -0.25 0.5\r\n+0.0625 .5 -.25\n0.3 0.3 0.3", 3)
    files["badline"] = (b"0.1 0.2 0.3\n0.1\t0.2 0.3\n", 3)
    ev = phantom_events(6000, 11)
    lines = []
    for k, (a, b, c) in enumerate(ev.astype(np.float64)):
        if k % 10 == 7:
            lines.append(f"{a:.9g} {b:.9g} {c:.9g}")
        elif k % 10 == 8:
            lines.append(f"{a:e} {b:e} {c:e}")
        else:
            lines.append(f"{a:.6g} {b:.6g} {c:.6g}")
    files["phantom"] = (("\n".join(lines) + "\n").encode(), 3)
    ev2 = phantom_events(1000, 12).astype(np.float64)
    files["exp2"] = (("\n".join(f"{a / 100:.9f}e2 {b / 100:.9f}e2 {c / 100:.9f}e2"
                                for a, b, c in ev2) + "\n").encode(), 3)
    with open(os.path.join(HERE, "convert", "warsaw_big.out"), "rb") as f:
        files["convert"] = (f.read(), 5)
    arrays = {"num_strings": np.array(nums, dtype="S"), "num_bits": bits, "num_ok": ok,
              "int_strings": np.array(ints, dtype="S"), "int_values": ivals}
    for name, (data, cols) in files.items():
        evs, n = load_file(ref, data, cols)
        arrays[f"file_{name}"] = np.frombuffer(data, np.uint8)
        arrays[f"events_{name}"] = evs
        arrays[f"n_{name}"] = np.int64(n)
        arrays[f"cols_{name}"] = np.int32(cols)
        print(f"{name}: {n} responses")
    np.savez_compressed(os.path.join(HERE, "text_responses.npz"), **arrays)
    print(f"{len(nums)} number strings, {int(ok.sum())} parse")


if __name__ == "__main__":
    main()
