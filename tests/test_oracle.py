"""CPU tests that pin the oracle before it is trusted as the checker.

* plan math (compute_m / compute_h / compute_h2 / shared_size) against the
  reference's own function.cpp compiled into oracle/_ref, and against closed
  forms;
* the FP64 DFT against numpy's FP64 FFT (golden fixtures), against the DFT
  definition evaluated by brute force in Python, and direct vs Bluestein;
* the FP32 restatement of the reference pipeline against the same vectors
  (within the FP32 tolerance), including its multi-tile global-stage FFT;
* mistakes planted in the oracle's use (wrong direction, transposed operand,
  dropped normalisation) are caught by the comparison the GPU tests use.
"""
import cmath
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import tol_nl2

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "dft2d_fixtures.npz")


def golden_cases():
    z = np.load(GOLDEN, allow_pickle=False)
    sizes = z["sizes"]
    return [(int(w), int(h), z[f"in_{i}"], z[f"fwd_{i}"], z[f"inv_{i}"]) for i, (w, h) in enumerate(sizes)]


def nl2(a, b):
    b = np.asarray(b, np.complex128)
    d = np.linalg.norm(np.asarray(a, np.complex128) - b)
    n = np.linalg.norm(b)
    return d / n if n > 0 else d


# ---------------------------------------------------------------- plan math

def test_compute_m_table():
    # padded lengths quoted for the configurations
    table = {640: 2048, 480: 1024, 1024: 1024, 4093: 8192, 2047: 4096, 6007: 16384, 997: 2048, 1: 1, 2: 2, 3: 4,
             5: 8, 17: 32, (1 << 29): (1 << 29)}
    for n, m in table.items():
        assert O.compute_m(n) == m, n
    for bad in (0, -1, (1 << 29) + 1):
        assert O.compute_m(bad) == -1


def test_compute_m_closed_form():
    for n in range(1, 5000):
        m = O.compute_m(n)
        if n & (n - 1) == 0:
            assert m == n
        else:
            assert m >= 2 * n - 2 and m & (m - 1) == 0 and m // 2 < 2 * n - 2


def ref_or_skip():
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return R


def test_compute_m_vs_reference():
    R = ref_or_skip()
    rng = np.random.default_rng(1)
    ns = list(range(-3, 4200)) + [int(v) for v in rng.integers(1, 1 << 29, 500)] + [(1 << 29), (1 << 29) + 1]
    for n in ns:
        assert O.compute_m(n) == R.ref_compute_m(n), n


@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 480, 640, 997, 2047, 4093, 6007, 46341, 100003])
@pytest.mark.parametrize("inverse", [0, 1])
def test_compute_h_h2_bitwise_vs_reference(n, inverse):
    R = ref_or_skip()
    m = O.compute_m(n)
    coef = (m / n) if inverse else 1.0
    mine = O.compute_h(n, inverse, coef)
    theirs = np.empty(n, np.complex128)
    R.ref_compute_h(n, inverse, coef, theirs.ctypes.data_as(O._dp))
    assert np.array_equal(mine, theirs)
    h2m = O.compute_h2(n, m, mine)
    h2r = np.empty(m, np.complex128)
    R.ref_compute_h2(n, m, theirs.ctypes.data_as(O._dp), h2r.ctypes.data_as(O._dp))
    assert np.array_equal(h2m, h2r)


@pytest.mark.parametrize("n", [3, 5, 17, 100, 997])
def test_compute_h_closed_form_and_h2_symmetry(n):
    m = O.compute_m(n)
    h = O.compute_h(n, 0, 1.0)
    l = np.arange(n)
    expect = np.exp(1j * np.pi * (l.astype(np.float64) ** 2) / n)
    assert np.abs(h - expect).max() < 1e-9
    h2 = O.compute_h2(n, m, h)
    # circulant embedding of an even kernel: h2[l] == h2[m-l], zeros in the gap
    assert np.array_equal(h2[1:], h2[1:][::-1])
    assert np.all(h2[n:m - n + 1] == 0)


def test_shared_size_vs_reference():
    R = ref_or_skip()
    for lds in (16384, 32768, 49152, 65536, 163840):
        for n in [1 << k for k in range(0, 16)]:
            assert O.lib().orc_shared_size(n, lds) == R.ref_shared_size(n, lds)


# ---------------------------------------------------------------- FP64 DFT

def test_fp64_against_golden():
    for (w, h, x, fwd, inv) in golden_cases():
        assert nl2(O.dft2d_64(x, False), fwd) < 1e-13, (w, h)
        assert nl2(O.dft2d_64(x, True), inv) < 1e-13, (w, h)


def test_constant_data_closed_form():
    # test_dft.cpp:391-400, w = 2, h = 3: X[k1,k2] = sum x[j2*2+j1] e^{-2pi i (j1 k1/2 + j2 k2/3)}
    x = np.array([1, 2, 3, 4, 5, 6], np.complex64).reshape(3, 2)
    s3 = math.sqrt(3.0)
    expect = np.array([21, -3, -6 + 2 * s3 * 1j, 0, -6 - 2 * s3 * 1j, 0])
    assert np.abs(O.dft2d_64(x, False).ravel() - expect).max() < 1e-12
    assert np.abs(O.dft2d_64(expect.reshape(3, 2), True) - x).max() < 1e-12
    # 1-D values quoted at test_dft.cpp:385-387
    y = O.dft1d_64(np.array([1, 2, 30]), False)
    assert np.abs(y - np.array([33, -15 + 14 * s3 * 1j, -15 - 14 * s3 * 1j])).max() < 1e-12
    assert np.abs(O.dft1d_64(np.array([1, 2]), False) - np.array([3, -1])).max() < 1e-15


def brute_dft2(x, inverse):
    h, w = x.shape
    sgn = 1 if inverse else -1
    out = np.zeros((h, w), complex)
    for k2 in range(h):
        for k1 in range(w):
            s = 0j
            for j2 in range(h):
                for j1 in range(w):
                    s += complex(x[j2, j1]) * cmath.exp(sgn * 2j * math.pi * (j1 * k1 / w + j2 * k2 / h))
            out[k2, k1] = s
    return out / (w * h) if inverse else out


@pytest.mark.parametrize("w,h", [(3, 5), (4, 2), (7, 1), (1, 6), (5, 5), (8, 3)])
def test_fp64_against_definition(w, h):
    rng = np.random.default_rng(w * 10 + h)
    x = (rng.standard_normal((h, w)) + 1j * rng.standard_normal((h, w))).astype(np.complex64)
    for inv in (False, True):
        assert nl2(O.dft2d_64(x, inv), brute_dft2(x.astype(complex), inv)) < 1e-13


def test_fp64_bluestein_against_direct():
    rng = np.random.default_rng(3)
    for n in list(range(65, 140)) + [257, 640, 997, 1000]:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        for inv in (False, True):
            a = O.dft1d_64(x, inv)
            b = O.dft1d_direct64(x, inv)
            assert nl2(a, b) < 1e-13, (n, inv)


def test_fp64_roundtrip_and_parseval():
    rng = np.random.default_rng(4)
    x = (rng.standard_normal((45, 77)) + 1j * rng.standard_normal((45, 77))).astype(np.complex64)
    y = O.dft2d_64(x, False, nthreads=4)
    assert nl2(O.dft2d_64(y, True), x) < 1e-14
    assert abs(np.sum(np.abs(y) ** 2) / (np.sum(np.abs(x.astype(complex)) ** 2) * x.size) - 1) < 1e-12


# ---------------------------------------------------------------- FP32 restatement

def test_ref32_against_golden():
    for (w, h, x, fwd, inv) in golden_cases():
        t = tol_nl2(w, h)
        assert nl2(O.dft2d_ref32(x, False), fwd) < t, (w, h)
        assert nl2(O.dft2d_ref32(x, True), inv) < t, (w, h)


@pytest.mark.parametrize("n,shared", [(1024, 128), (4096, 256), (16384, 8192), (256, 8192), (2, 128)])
def test_fft32_shared_and_global_stages(n, shared):
    # fft.cpp:293-329: n > shared runs bit reverse + LDS tiles + global stages
    rng = np.random.default_rng(n + shared)
    x = (rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))).astype(np.complex64)
    f = O.fft32(x.ravel(), n, False, shared).reshape(3, n)
    assert nl2(f, np.fft.fft(x.astype(complex), axis=1)) < 1e-6
    i = O.fft32(x.ravel(), n, True, shared).reshape(3, n)
    assert nl2(i, np.fft.ifft(x.astype(complex), axis=1)) < 1e-6


@pytest.mark.parametrize("n1,n2", [(640, 480), (17, 2), (997, 6)])
def test_diagonals32_against_fp64(n1, n2):
    d1f, d1i, d2f, d2i = O.diagonals32(n1, n2)
    for n, df, di in ((n1, d1f, d1i), (n2, d2f, d2i)):
        m = O.compute_m(n)
        hf = O.compute_h2(n, m, O.compute_h(n, 0, 1.0))
        hi = O.compute_h2(n, m, O.compute_h(n, 1, m / n))
        assert nl2(df, np.fft.fft(hf)) < 2e-6 * math.log2(m) + 1e-7
        assert nl2(di, np.fft.ifft(hi)) < 2e-6 * math.log2(m) + 1e-7


def test_ref32_threaded_equals_serial():
    rng = np.random.default_rng(8)
    x = (rng.uniform(-1, 1, (61, 90)) + 1j * rng.uniform(-1, 1, (61, 90))).astype(np.complex64)
    assert np.array_equal(O.dft2d_ref32(x, False, nthreads=1), O.dft2d_ref32(x, False, nthreads=7))


# ---------------------------------------------------------------- planted mistakes

def test_planted_mistakes_are_caught():
    rng = np.random.default_rng(12)
    w, h = 37, 24
    x = (rng.uniform(-1, 1, (h, w)) + 1j * rng.uniform(-1, 1, (h, w))).astype(np.complex64)
    ref = O.dft2d_64(x, False)
    t = tol_nl2(w, h)
    assert O.error_metrics(O.dft2d_ref32(x, False), ref)["nl2"] < t
    # wrong direction
    assert O.error_metrics(O.dft2d_ref32(x, True) * (w * h), ref)["nl2"] > 100 * t
    # transposed operand
    xt = np.ascontiguousarray(x.T.copy()[:h * w].reshape(h, w))
    assert O.error_metrics(O.dft2d_ref32(xt, False), ref)["nl2"] > 100 * t
    # dropped inverse normalisation
    inv = O.dft2d_64(x, True)
    assert O.error_metrics((O.dft2d_ref32(x, True) * w).astype(np.complex64), inv)["nl2"] > 100 * t
    # one axis only
    assert O.error_metrics(np.fft.fft(x, axis=1).astype(np.complex64), ref)["nl2"] > 100 * t


def test_fill_deterministic_and_ranges():
    a = O.fill(1, "uniform", (100, 100))
    b = O.fill(1, "uniform", (100, 100))
    assert np.array_equal(a, b)
    assert a.real.min() >= -1 and a.real.max() < 1
    n = O.fill(2, "normal", (300, 300))
    assert abs(n.real.mean()) < 0.02 and abs(n.real.std() - 1) < 0.02


# ---- image path (src/gpu/dft/compute.cpp:91-200) ----------------------------

def test_image_luminance_weights():
    # dft_copy_input.comp:37-40: pure primaries give the weights, white gives their sum
    px = np.zeros((1, 4, 4), np.float32)
    px[0, 0, :3] = (1, 0, 0)
    px[0, 1, :3] = (0, 1, 0)
    px[0, 2, :3] = (0, 0, 1)
    px[0, 3, :3] = (1, 1, 1)
    lum = O.image_luminance(px, 0, 0, 4, 1)[0]
    assert np.allclose(lum, [0.2126, 0.7152, 0.0722, 1.0], rtol=0, atol=1e-7)
    # UNORM bytes: 255 is 1.0, alpha is ignored
    p8 = np.zeros((1, 2, 4), np.uint8)
    p8[0, 0] = (255, 255, 255, 0)
    p8[0, 1] = (51, 0, 0, 255)
    lum8 = O.image_luminance(p8, 0, 0, 2, 1)[0]
    assert abs(lum8[0] - 1.0) < 1e-7 and abs(lum8[1] - 0.2126 * 0.2) < 1e-7


def test_image_luminance_rectangle_offset():
    rng = np.random.default_rng(4)
    px = rng.uniform(0, 1, (9, 11, 4)).astype(np.float32)
    lum = O.image_luminance(px, 3, 2, 5, 4)
    sub = px[2:6, 3:8].astype(np.float64)
    expect = 0.2126 * sub[..., 0] + 0.7152 * sub[..., 1] + 0.0722 * sub[..., 2]
    assert np.abs(lum - expect).max() < 1e-6


def test_image_dft_closed_forms():
    w, h = 6, 5
    # constant luminance l: only the DC term, |X(0,0)|/(w*h) = l
    px = np.zeros((h, w, 4), np.float32)
    px[..., :3] = 0.5
    out = O.image_dft_64(px, 0, 0, w, h)
    assert abs(out[0, 0] - 0.5) < 1e-7
    out[0, 0] = 0
    assert np.abs(out).max() < 1e-12
    # an impulse of luminance l has a flat spectrum l/(w*h)
    px = np.zeros((h, w, 4), np.float32)
    px[0, 0, :3] = 1.0
    out = O.image_dft_64(px, 0, 0, w, h)
    assert np.abs(out - 1.0 / (w * h)).max() < 1e-7
