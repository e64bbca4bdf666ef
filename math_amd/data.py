"""Text data files of the reference's DFT test (src/gpu/dft/test/test_data.cpp:36-125),
the Python side of include/mdft_data.hpp: load_data, save_data and
generate_random_data write and read the same bytes as the C++ header.

Random data: a PCG-XSH-RR engine (src/com/random/pcg.h:53-63) seeded with four
words of std::seed_seq({seed}) (pcg.cpp:45-80; the seed_seq algorithm of the C++
standard, [rand.util.seedseq]), drawn through libstdc++'s
uniform_real_distribution(-1, 1) (generate_canonical with one 32-bit draw per
float, two per double, then u * (b - a) + a in the value type).
"""
import math

import numpy as np

MASK32 = 0xFFFFFFFF
MASK64 = 0xFFFFFFFFFFFFFFFF


class DataError(RuntimeError):
    pass


def seed_seq_generate(values, n):
    """std::seed_seq(values).generate() of n 32-bit words ([rand.util.seedseq])."""
    v = [x & MASK32 for x in values]
    s = len(v)
    b = [0x8B8B8B8B] * n
    if n == 0:
        return b
    t = 11 if n >= 623 else 7 if n >= 68 else 5 if n >= 39 else 3 if n >= 7 else (n - 1) // 2
    p = (n - t) // 2
    q = p + t
    m = max(s + 1, n)

    def T(x):
        return x ^ (x >> 27)

    for k in range(m):
        r1 = (1664525 * T(b[k % n] ^ b[(k + p) % n] ^ b[(k - 1) % n])) & MASK32
        if k == 0:
            r2 = r1 + s
        elif k <= s:
            r2 = r1 + k % n + v[k - 1]
        else:
            r2 = r1 + k % n
        r2 &= MASK32
        b[(k + p) % n] = (b[(k + p) % n] + r1) & MASK32
        b[(k + q) % n] = (b[(k + q) % n] + r2) & MASK32
        b[k % n] = r2
    for k in range(m, m + n):
        r3 = (1566083941 * T((b[k % n] + b[(k + p) % n] + b[(k - 1) % n]) & MASK32)) & MASK32
        r4 = (r3 - k % n) & MASK32
        b[(k + p) % n] ^= r3
        b[(k + q) % n] ^= r4
        b[k % n] = r4
    return b


class Pcg:
    """PCG-XSH-RR 64/32 with the reference's seeding (pcg.cpp:45-80)."""

    def __init__(self, seed):
        w = seed_seq_generate([seed & MASK32], 4)
        self.state = w[0] | (w[1] << 32)
        self.inc = w[2] | (w[3] << 32) | 1

    def __call__(self):
        x = self.state
        self.state = (x * 6364136223846793005 + self.inc) & MASK64
        s = (((x >> 18) ^ x) >> 27) & MASK32
        r = x >> 59
        return ((s >> r) | (s << ((32 - r) & 31))) & MASK32


def uniform_float(engine, a=-1.0, b=1.0):
    """libstdc++ uniform_real_distribution<float>(a, b)(engine)."""
    f32 = np.float32
    u = f32(engine()) / f32(4294967296.0)
    if u >= f32(1):
        u = np.nextafter(f32(1), f32(0))
    return f32(u * (f32(b) - f32(a))) + f32(a)


def uniform_double(engine, a=-1.0, b=1.0):
    """libstdc++ uniform_real_distribution<double>(a, b)(engine)."""
    lo = float(engine())
    hi = float(engine())
    u = (lo + hi * 4294967296.0) / 18446744073709551616.0
    if u >= 1.0:
        u = math.nextafter(1.0, 0.0)
    return u * (b - a) + a


def _fmt(dtype):
    return "{:+.9e}" if dtype == np.float32 else "{:+.17e}"


def save_data(path, x):
    """test_data.cpp:83-97: one "re im" line per value, no header."""
    x = np.asarray(x)
    dtype = np.float32 if x.dtype == np.complex64 else np.float64
    f = _fmt(dtype)
    with open(path, "w") as out:
        for c in x.ravel():
            out.write(f.format(float(c.real)) + " " + f.format(float(c.imag)) + "\n")


def load_data(path, dtype=np.complex64):
    """test_data.cpp:36-81: returns (n1, n2, data) with the reference's errors."""
    with open(path) as f:
        tokens = f.read().split()
    try:
        n1, n2 = int(tokens[0]), int(tokens[1])
    except (IndexError, ValueError):
        raise DataError("Data dimensions read error")
    if n1 < 1 or n2 < 1:
        raise DataError("Dimensions must be positive numbers")
    if n1 > 1_000_000_000 or n2 > 1_000_000_000:
        raise DataError("Dimensions are too big")
    count = n1 * n2
    real = np.float32 if dtype == np.complex64 else np.float64
    vals = tokens[2:2 + 2 * count]
    out = np.empty(count, dtype)
    for i in range(count):
        try:
            re, im = real(vals[2 * i]), real(vals[2 * i + 1])
        except (IndexError, ValueError):
            raise DataError(f"Error reading number {i}")
        out[i] = complex(re, im)
    return n1, n2, out


def generate_random_data(path, n1, n2, dtype=np.complex64):
    """test_data.cpp:99-125; the reference's seed is n2 (its 64-bit
    (n1 << 32) + n2 is narrowed to the engine's 32-bit seed)."""
    if n1 < 1 or n2 < 1:
        raise DataError(f"Wrong size {n1} {n2}")
    eng = Pcg(((n1 << 32) + n2) & MASK32)
    draw = uniform_float if dtype == np.complex64 else uniform_double
    f = _fmt(np.float32 if dtype == np.complex64 else np.float64)
    with open(path, "w") as out:
        out.write(f"{n1} {n2}\n")
        for _ in range(n1 * n2):
            re = draw(eng)
            im = draw(eng)
            out.write(f.format(float(re)) + " " + f.format(float(im)) + "\n")
