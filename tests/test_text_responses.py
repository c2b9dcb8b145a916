"""Text response files read as the reference's fast text reader reads them
(util::text_parser, src/util/text_parser.h:29-175; Strip::Reconstruction::
fast_load_txt_events, src/2d/strip/reconstruction.h:169-196; the 3D hybrid
reader, src/3d/hybrid/reconstruction.h:165-188), on the CPU.

The expected values are the reference's own: tests/golden/text_responses.npz
holds what the reader, compiled from its headers, returned for each string
and file (written by tests/golden/make_text_golden.py).  The comparisons are
bit for bit.
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "text_responses.npz"))
FILES = ["quirks", "badline", "phantom", "exp2", "convert"]


def test_number_strings_bitwise(pkg):
    strs, bits, ok = GOLD["num_strings"], GOLD["num_bits"], GOLD["num_ok"]
    assert len(strs) > 2000 and 0 < ok.sum() < len(strs)
    for s, b, good in zip(strs, bits, ok):
        if good:
            v, _ = pkg.parse_text_floats(s, 1)
            assert v.view(np.uint32)[0] == b, (s, v[0], np.uint32(b).view(np.float32))
        else:
            with pytest.raises(pkg.PetstripError):
                pkg.parse_text_floats(s, 1)


def test_fixtures_tell_the_reader_apart():
    """The fixture strings separate the reference's reader from a correctly
    rounded one and from the same formula without the fused multiply-add, so
    a reader that took either form would fail the bitwise test above."""
    strs, bits, ok = GOLD["num_strings"], GOLD["num_bits"], GOLD["num_ok"]
    n_strtod = n_unfused = 0
    for s, b, good in zip(strs, bits, ok):
        t = s.decode()
        if not good or any(c in t for c in "eE ") or t.strip() != t:
            continue
        try:
            if np.float32(float(t)).view(np.uint32) != b:
                n_strtod += 1
        except ValueError:
            continue
        neg = t.startswith("-")
        body = t.lstrip("+-")
        whole, _, frac = body.partition(".")
        if not frac or len(whole) > 9 or len(frac) > 9:
            continue
        p = np.float32(1.0)
        for _ in range(len(frac)):
            p = np.float32(np.float32(0.1) * p)
        v = np.float32(np.float32(int(whole or "0")) +
                       np.float32(np.float32(int(frac)) * p))
        v = np.float32(v * np.float32(-1.0 if neg else 1.0))
        if v.view(np.uint32) != b:
            n_unfused += 1
    assert n_strtod > 300 and n_unfused > 5, (n_strtod, n_unfused)


def test_int_fields(pkg, tmp_path):
    """The LOR fields of the 5-field form (operator>><int>, then int16 in
    the hybrid's LOR), read from lines whose float fields parse."""
    for s, v in zip(GOLD["int_strings"], GOLD["int_values"]):
        t = s.decode()
        if not t.lstrip("+-")[:1].isdigit():
            continue
        f = tmp_path / "l.txt"
        f.write_text(f"{t} 3 0.25 -0.125 0.5\n")
        ev, lors = pkg.read_text_responses(f, 5, with_lors=True)
        want = np.int32(v).astype(np.int16)  # the 16-bit LOR field
        if "." in t:
            continue  # "12.5" as an int stops at the period; the next field fails
        assert lors[0, 0] == want and lors[0, 1] == 3, (t, lors[0], want)
        assert np.array_equal(ev[0], pkg.parse_text_floats("0.25 -0.125 0.5", 3)[0])


@pytest.mark.parametrize("name", FILES)
def test_files_bitwise(pkg, tmp_path, name):
    data = GOLD[f"file_{name}"].tobytes()
    cols = int(GOLD[f"cols_{name}"])
    n = int(GOLD[f"n_{name}"])
    path = tmp_path / f"{name}.txt"
    path.write_bytes(data)
    if n < 0:
        with pytest.raises(pkg.PetstripError, match="error line"):
            pkg.read_text_responses(path, cols)
        return
    ev = pkg.read_text_responses(path, cols)
    want = GOLD[f"events_{name}"]
    assert ev.shape == (n, 3)
    assert np.array_equal(ev.view(np.uint32), want.view(np.uint32))


def test_last_line_and_long_lines(pkg, tmp_path):
    """A last line without '\\n' is not read; reading stops at a line of
    16384 bytes or more (the reader's buffer), keeping the lines before."""
    p = tmp_path / "a.txt"
    p.write_bytes(b"0.5 0.5 0.5\n0.25 0.25 0.25")
    assert pkg.read_text_responses(p).tolist() == [[0.5, 0.5, 0.5]]
    long_line = b"0.125 0.125 0.125" + b" " * (16384 - 17)
    p.write_bytes(b"0.5 0.5 0.5\n" + long_line[:-1] + b"\n0.25 0.25 0.25\n")
    assert pkg.read_text_responses(p).shape == (3, 3)
    p.write_bytes(b"0.5 0.5 0.5\n" + long_line + b"\n0.25 0.25 0.25\n")
    assert pkg.read_text_responses(p).tolist() == [[0.5, 0.5, 0.5]]
    p.write_bytes(b"")
    assert pkg.read_text_responses(p).shape == (0, 3)
    with pytest.raises(pkg.PetstripError, match="cannot read"):
        pkg.read_text_responses(tmp_path / "missing.txt")
    with pytest.raises(pkg.PetstripError):
        pkg.read_text_responses(p, 4)


def test_exponent_quirk(pkg):
    """e > 0 multiplies by 10 * 0.1f^(e - 1), so "1.5e3" reads as 0.15
    (text_parser.h:83-91, 168-169); e < 0 by 0.1f^-e."""
    v, used = pkg.parse_text_floats("1.5e3 2e-2 -4E+1", 3)
    p1 = np.float32(10.0) * np.float32(np.float32(0.1) * np.float32(0.1))
    assert v[0] == np.float32(np.float32(1.5) * p1)
    assert v[1] == np.float32(2.0) * np.float32(np.float32(0.1) * np.float32(0.1))
    assert v[2] == np.float32(-4.0) * np.float32(10.0)
    assert used == len("1.5e3 2e-2 -4E+1")
