"""The reference's DFT test data files (src/gpu/dft/test/test_data.cpp:36-125):
the Python reader/writer (math_amd/data.py) and the C++ header
(include/mdft_data.hpp) write the same bytes and read each other's files.

Pinned by the PCG known answers of src/com/random/test/test_pcg.cpp:191-192,
by the C++ header compiled here against this toolchain's standard library
(std::seed_seq, std::uniform_real_distribution), and by exact round trips.
The reference's own generated values depend on its standard library's
uniform_real_distribution: with the same library they match byte for byte;
across libraries that part is parity unpinned."""
import os
import subprocess

import numpy as np
import pytest

from math_amd import data as D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def tool(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("tool") / "data_tool")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "data_tool.cpp"), "-o", exe], check=True)
    return exe


def test_pcg_known_answers():
    # test_pcg.cpp:191-192: the 1000th output of PCG(0) and PCG(1000)
    for seed, expect in ((0, 1557370411), (1000, 2243789472)):
        e = D.Pcg(seed)
        v = None
        for _ in range(1000):
            v = e()
        assert v == expect


def test_pcg_matches_cpp(tool):
    for seed in (0, 1, 97, 4096, 0xFFFFFFFF):
        out = subprocess.run([tool, "pcg", str(seed), "77"], capture_output=True, text=True, check=True).stdout
        e = D.Pcg(seed)
        v = [e() for _ in range(77)][-1]
        assert int(out) == v


@pytest.mark.parametrize("n1,n2,dtype", [(1, 1, np.complex64), (37, 23, np.complex64), (100, 3, np.complex64),
                                         (5, 7, np.complex128), (1, 64, np.complex128)])
def test_generated_files_identical(tool, tmp_path, n1, n2, dtype):
    a, b = str(tmp_path / "a.txt"), str(tmp_path / "b.txt")
    args = [tool, "gen", a, str(n1), str(n2)] + (["d"] if dtype == np.complex128 else [])
    subprocess.run(args, check=True)
    D.generate_random_data(b, n1, n2, dtype)
    assert open(a, "rb").read() == open(b, "rb").read()
    m1, m2, x = D.load_data(b, dtype)
    assert (m1, m2) == (n1, n2) and x.size == n1 * n2
    assert np.all(np.abs(x.real) <= 1) and np.all(np.abs(x.imag) <= 1)


def test_seed_is_the_height(tmp_path):
    # the 64-bit (n1 << 32) + n2 is narrowed to a 32-bit seed: n2
    a, b = str(tmp_path / "a.txt"), str(tmp_path / "b.txt")
    D.generate_random_data(a, 3, 11)
    D.generate_random_data(b, 5, 11)
    la = open(a).read().splitlines()[1:]
    lb = open(b).read().splitlines()[1:]
    assert la == lb[:len(la)]


def test_save_load_round_trip_exact(tool, tmp_path):
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(500) * 10.0 ** rng.integers(-30, 30, 500) + 1j * rng.standard_normal(500)).astype(
        np.complex64)
    x[0] = 0
    x[1] = complex(np.finfo(np.float32).max, -np.finfo(np.float32).tiny)
    p = str(tmp_path / "x.txt")
    D.save_data(p, x)
    body = open(p).read()
    with open(str(tmp_path / "h.txt"), "w") as f:
        f.write("25 20\n" + body)
    n1, n2, y = D.load_data(str(tmp_path / "h.txt"))
    assert (n1, n2) == (25, 20)
    assert np.array_equal(x, y)
    # the C++ reader and writer reproduce the file
    out = str(tmp_path / "c.txt")
    subprocess.run([tool, "copy", str(tmp_path / "h.txt"), out], check=True, capture_output=True)
    assert open(out).read() == body


def test_number_format():
    # scientific, max_digits10 digits after the point, explicit sign (test_data.cpp:86-90)
    assert D._fmt(np.float32).format(0.5) == "+5.000000000e-01"
    assert D._fmt(np.float64).format(-0.1) == "-1.00000000000000006e-01"


@pytest.mark.parametrize("text,msg", [("x 3\n", "Data dimensions read error"), ("0 3\n", "positive"),
                                      ("2000000000 1\n", "too big"), ("2 1\n+1 +2\n+3\n", "Error reading number 1")])
def test_load_errors(tool, tmp_path, text, msg):
    p = str(tmp_path / "bad.txt")
    open(p, "w").write(text)
    with pytest.raises(D.DataError, match=msg):
        D.load_data(p)
    r = subprocess.run([tool, "copy", p, str(tmp_path / "o.txt")], capture_output=True, text=True)
    assert r.returncode == 2 and msg in r.stdout
