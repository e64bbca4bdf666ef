"""GPU parity of the HIP path against the CPU oracle (FP64 DFT and the FP32
restatement of the reference pipeline), through the C ABI.

Mirrors src/gpu/dft/test/test_dft.cpp: the constant-data test (:383-407), the
random-data test at small sizes and at 6007x997 (:409-462), re-planning at
random sizes including 1x1, 1xk and kx1 before the timed run (:138-148), and
the reference's pass criterion (relative L1 <= 1e-4, :59,63-93), plus the
tighter normalised-L2 bound 2e-6*log2(padded N) against the FP64 oracle.
"""
import math
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import REF_REL_L1, check_against_fp64, rand_c64, tol_nl2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mdft():
    from math_amd import mdft as M

    M.init(0)
    return M


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_exec(M, x, inverse):
    h, w = x.shape
    d = np.ascontiguousarray(x.copy())
    p = M.Plan(w, h)
    p.exec(inverse, d)
    p.close()
    return d


def test_constant_data(mdft):
    # test_dft.cpp:391-400: {1..6} as n1 = 2 (width) x n2 = 3 (height)
    x = np.array([1, 2, 3, 4, 5, 6], np.complex64).reshape(3, 2)
    s3 = math.sqrt(3.0)
    expect = np.array([21, -3, -6 + 2 * s3 * 1j, 0, -6 - 2 * s3 * 1j, 0], np.complex128)
    cv = mdft.create_compute_vector()
    cv.create_buffers(2, 3)
    d = x.copy().ravel()
    cv.exec(False, d)
    assert np.abs(d - expect).max() < 1e-5, d
    cv.exec(True, d)
    assert np.abs(d - x.ravel()).max() < 1e-5, d


def test_one_dimensional_kats(mdft):
    # the Mathematica values quoted at test_dft.cpp:385-387
    d = np.array([1, 2, 30], np.complex64)
    p = mdft.Plan(3, 1)
    p.exec(False, d)
    expect = np.array([33, -15 + 14 * math.sqrt(3) * 1j, -15 - 14 * math.sqrt(3) * 1j])
    assert np.abs(d - expect).max() < 1e-4, d
    d = np.array([1, 2], np.complex64)
    mdft.Plan(1, 2).exec(False, d)
    assert np.abs(d - np.array([3, -1])).max() < 1e-6, d


SMALL_SIZES = [(1, 1), (1, 7), (7, 1), (2, 2), (3, 5), (16, 16), (17, 13), (31, 64), (64, 31), (100, 1), (1, 100),
               (37, 61), (97, 89), (100, 100), (128, 3), (5, 256)]


@pytest.mark.parametrize("w,h", SMALL_SIZES)
def test_small_sizes(mdft, w, h):
    rng = np.random.default_rng(w * 1000 + h)
    x = rand_c64(rng, (h, w))
    log(f"small {w}x{h}")
    for inv in (False, True):
        got = run_exec(mdft, x, inv)
        check_against_fp64(got, x, inv, "small")
        ref32 = O.dft2d_ref32(x, inv, nthreads=8)
        m = O.error_metrics(got, ref32.astype(np.complex128))
        assert m["nl2"] <= 2 * tol_nl2(w, h), m


def test_random_sweep_1_to_100(mdft):
    # test_dft.cpp:437-449: 90 % of sizes uniform in [1, 100]^2
    rng = np.random.default_rng(2024)
    for _ in range(40):
        w, h = int(rng.integers(1, 101)), int(rng.integers(1, 101))
        x = rand_c64(rng, (h, w))
        inv = bool(rng.integers(0, 2))
        got = run_exec(mdft, x, inv)
        check_against_fp64(got, x, inv, "sweep")


def test_random_sweep_to_4096(mdft):
    # test_dft.cpp:437-449 draws 10 % of its sizes up to 1000; here log-uniform sides in [100, 4096]
    # (every shape family on both axes: powers of two, 3 * 2^k and 2^k convolutions, direct 3 * 2^k),
    # forward and inverse, through the device entry point
    import torch

    rng = np.random.default_rng(3000)
    for _ in range(160):
        w, h = (int(np.exp(rng.uniform(np.log(100), np.log(4097)))) for _ in range(2))
        x = rand_c64(rng, (h, w))
        inv = bool(rng.integers(0, 2))
        src = torch.from_numpy(x.ravel().copy()).cuda()
        dst = torch.empty_like(src)
        p = mdft.Plan(w, h)
        p.exec_device(inv, src.data_ptr(), dst.data_ptr(), torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        p.close()
        log(f"sweep {w}x{h} inv={inv}")
        check_against_fp64(dst.cpu().numpy().reshape(h, w), x, inv, "sweep to 4096")


def test_all_line_lengths_1_to_300(mdft):
    # every line length of one axis: exercises every padded length up to 1024 and both paths
    rng = np.random.default_rng(7)
    for n in range(1, 301):
        x = rand_c64(rng, (3, n))
        got = run_exec(mdft, x, False)
        check_against_fp64(got, x, False, "rows")
        xt = rand_c64(rng, (n, 2))
        got = run_exec(mdft, xt, True)
        check_against_fp64(got, xt, True, "cols")


# 1025-1536 and 2049-3072 (and 2500) run 3 * 2^k convolutions along both axes, 3073 the next power of two;
# 384, 768, 1536, 3072 and 6144 are 3 * 2^k themselves (direct FFTs)
# 8193 is the longest Bluestein line one tile holds (16384-point convolution), 16383 the shortest above it
@pytest.mark.parametrize("n", [384, 511, 512, 513, 768, 1000, 1023, 1024, 1025, 1536, 2047, 2048, 2049, 2500, 3072,
                               3073, 4093, 4096, 4097, 6144, 8191, 8192, 8193, 8194, 16383, 16384])
def test_long_lines(mdft, n):
    rng = np.random.default_rng(n)
    x = rand_c64(rng, (2, n))
    for inv in (False, True):
        check_against_fp64(run_exec(mdft, x, inv), x, inv, "long rows")
    xt = rand_c64(rng, (n, 3))
    for inv in (False, True):
        check_against_fp64(run_exec(mdft, xt, inv), xt, inv, "long cols")


def test_config_640x480_fwd_inv(mdft):
    rng = np.random.default_rng(0x6D696C6C + 0)
    x = rand_c64(rng, (480, 640), "uniform")
    log("640x480")
    y = run_exec(mdft, x, False)
    m = check_against_fp64(y, x, False, "C1 fwd")
    ref32 = O.dft2d_ref32(x, False, nthreads=8)
    m32 = O.error_metrics(y, ref32.astype(np.complex128))
    assert m32["nl2"] <= tol_nl2(640, 480), m32
    z = run_exec(mdft, y, True)
    check_against_fp64(z, y, True, "C1 inv")
    rt = O.error_metrics(z, x.astype(np.complex128))
    assert rt["nl2"] <= tol_nl2(640, 480), rt


def check_against_ref32(got, x, inverse, label):
    """Agreement with the FP32 restatement of the reference GPU pipeline
    (oracle orc_dft2d_ref32, src/gpu/dft/dft.cpp:326-347) within
    2e-6*log2(padded N)."""
    h, w = x.shape
    ref32 = O.dft2d_ref32(x, inverse, nthreads=8)
    m = O.error_metrics(got.reshape(h, w), ref32.astype(np.complex128))
    assert m["nl2"] <= tol_nl2(w, h), f"{label} {w}x{h} inv={inverse}: nl2 vs FP32 pipeline {m['nl2']:.3e}"
    return m


def test_config_1024x1024_round_trip(mdft):
    rng = np.random.default_rng(0x6D696C6C + 1)
    x = rand_c64(rng, (1024, 1024), "normal")
    log("1024x1024")
    y = run_exec(mdft, x, False)
    check_against_fp64(y, x, False, "C2 fwd")
    check_against_ref32(y, x, False, "C2 fwd")
    z = run_exec(mdft, y, True)
    check_against_ref32(z, y, True, "C2 inv")
    rt = O.error_metrics(z, x.astype(np.complex128))
    assert rt["nl2"] <= tol_nl2(1024, 1024), rt
    # Parseval: ||X||^2 = w*h*||x||^2
    px = np.sum(np.abs(x.astype(np.complex128)) ** 2) * 1024 * 1024
    py = np.sum(np.abs(y.astype(np.complex128)) ** 2)
    assert abs(py / px - 1) < 1e-5


def test_config_4093x2047(mdft):
    rng = np.random.default_rng(0x6D696C6C + 2)
    x = rand_c64(rng, (2047, 4093), "uniform")
    log("4093x2047")
    for inv in (False, True):
        got = run_exec(mdft, x, inv)
        check_against_fp64(got, x, inv, "C3")
        check_against_ref32(got, x, inv, "C3")


def test_reference_large_6007x997(mdft):
    # test_dft.cpp:454-458 (padded 16384 x 2048)
    rng = np.random.default_rng(6007)
    x = rand_c64(rng, (997, 6007), "uniform")
    log("6007x997")
    for inv in (False, True):
        got = run_exec(mdft, x, inv)
        check_against_fp64(got, x, inv, "6007x997")
        check_against_ref32(got, x, inv, "6007x997")


def test_replanning(mdft):
    # test_dft.cpp:138-148: re-plan at random sizes, then the real one
    rng = np.random.default_rng(11)
    cv = mdft.create_compute_vector()
    for (w, h) in [(int(rng.integers(1, 3001)), int(rng.integers(1, 3001))), (1, 1),
                   (int(rng.integers(1, 3001)), int(rng.integers(1, 3001))), (1, int(rng.integers(1, 3001))),
                   (int(rng.integers(1, 3001)), 1)]:
        cv.create_buffers(w, h)
        x = rand_c64(rng, (h, w))
        d = x.ravel().copy()
        cv.exec(False, d)
        check_against_fp64(d, x, False, "replan")
    cv.create_buffers(90, 45)
    x = rand_c64(rng, (45, 90))
    d = x.ravel().copy()
    cv.exec(True, d)
    check_against_fp64(d, x, True, "replan final")


def test_errors(mdft):
    cv = mdft.create_compute_vector()
    with pytest.raises(mdft.ErrorException, match="No DFT buffers"):
        cv.exec(False, np.zeros(6, np.complex64))
    cv.create_buffers(2, 3)
    with pytest.raises(mdft.ErrorException, match="Wrong DFT buffer size"):
        cv.exec(False, np.zeros(5, np.complex64))
    with pytest.raises(mdft.ErrorException, match="Error size 0 for compute m"):
        cv.create_buffers(0, 3)
    with pytest.raises(mdft.ErrorException, match="for compute m"):
        mdft.Plan(3, (1 << 29) + 1)


def test_exec_device_in_and_out_of_place(mdft):
    import torch

    rng = np.random.default_rng(5)
    for (w, h) in [(640, 480), (33, 1), (1, 45), (1, 1), (257, 129)]:
        x = rand_c64(rng, (h, w))
        src = torch.from_numpy(x.copy()).to("cuda")
        dst = torch.empty_like(src)
        p = mdft.Plan(w, h)
        s = torch.cuda.current_stream().cuda_stream
        p.exec_device(False, src.data_ptr(), dst.data_ptr(), s)
        torch.cuda.synchronize()
        out = dst.cpu().numpy()
        check_against_fp64(out, x, False, "device oop")
        assert np.array_equal(src.cpu().numpy(), x), "source modified by an out-of-place transform"
        p.exec_device(True, src.data_ptr(), src.data_ptr(), s)
        torch.cuda.synchronize()
        check_against_fp64(src.cpu().numpy(), x, True, "device in place")


def test_host_and_device_paths_bitwise_equal(mdft):
    import torch

    rng = np.random.default_rng(9)
    x = rand_c64(rng, (77, 130))
    a = run_exec(mdft, x, False)
    t = torch.from_numpy(x.copy()).cuda()
    p = mdft.Plan(130, 77)
    p.exec_device(False, t.data_ptr(), t.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert np.array_equal(a, t.cpu().numpy())


@pytest.mark.parametrize("w,h", [(16385, 2), (2, 16385), (32768, 3), (3, 32768), (65536, 1), (1, 20000),
                                 (40000, 1), (9000, 5)])
def test_lines_longer_than_one_tile(mdft, w, h):
    # padded length above 16384: the four-step path through plan-owned scratch
    rng = np.random.default_rng(w + 7 * h)
    x = rand_c64(rng, (h, w))
    log(f"long line {w}x{h}")
    for inv in (False, True):
        check_against_fp64(run_exec(mdft, x, inv), x, inv, "four-step")


def test_longest_line_limit_message(mdft):
    # the longest side is compute_m's (function.cpp:34-38); the lines up to it run in
    # tests/test_gpu_long_lines.py
    with pytest.raises(mdft.ErrorException, match="Error size 536870913 for compute m"):
        mdft.Plan((1 << 29) + 1, 1)
    with pytest.raises(mdft.ErrorException, match="exceeds the longest supported line"):
        mdft.create_fft(1, 1 << 30)
