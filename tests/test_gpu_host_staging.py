"""Host-vector exec of large images: transfers of 16 MiB and more go through a
pinned two-chunk ring (4 MiB chunks) instead of the runtime's pageable copies.
The host-vector result must equal the device-resident call's bitwise (same
kernels, only the copies differ), for sizes that are not a multiple of the
chunk, for the 2-D plan (mdft_exec, compute.cpp:257-283) and the 1-D FFT
(mdft_fft_exec, fft.cpp:264-290); numpy's FP64 FFT bounds the error."""
import math

import numpy as np
import pytest

from tests.helpers import rand_c64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mdft():
    from math_amd import mdft as M

    M.init(0)
    return M


def padded(n):
    """Padded FFT length of a side (src/gpu/dft/function.cpp:32-45)."""
    return n if n & (n - 1) == 0 else 1 << (2 * n - 3).bit_length()


def nl2(got, ref):
    return float(np.linalg.norm(got.astype(np.complex128) - ref) / np.linalg.norm(ref))


def device_result(exec_device, x):
    import torch

    src = torch.from_numpy(x.copy().ravel()).cuda()
    dst = torch.empty_like(src)
    exec_device(src.data_ptr(), dst.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return dst.cpu().numpy()


@pytest.mark.parametrize("w,h", [(2100, 1100), (1031, 2500)])
def test_plan_host_vector_equals_device(mdft, w, h):
    # 18.5 MB and 20.6 MB: several chunks and a ragged last one
    assert 8 * w * h >= 16 << 20 and (8 * w * h) % (4 << 20) != 0
    rng = np.random.default_rng(w + h)
    x = rand_c64(rng, (h, w))
    p = mdft.Plan(w, h)
    for inv in (False, True):
        d = x.copy().ravel()
        p.exec(inv, d)
        dev = device_result(lambda s, t, st: p.exec_device(inv, s, t, st), x)
        assert np.array_equal(d, dev), (w, h, inv)
        ref = np.fft.ifft2(x.astype(np.complex128)) if inv else np.fft.fft2(x.astype(np.complex128))
        assert nl2(d.reshape(h, w), ref) <= 2e-6 * math.log2(padded(w) * padded(h)), (w, h, inv)
    p.close()


def test_fft1d_host_vector_equals_device(mdft):
    count, n = 1100, 4096  # 36 MB
    rng = np.random.default_rng(5)
    x = rand_c64(rng, (count, n))
    f = mdft.Fft(count, n)
    for inv in (False, True):
        d = x.copy().ravel()
        f.run_for_data(inv, d)
        dev = device_result(lambda s, t, st: f.exec_device(inv, s, t, st), x)
        assert np.array_equal(d, dev), inv
        xr = x.astype(np.complex128)
        ref = np.fft.ifft(xr, axis=1) if inv else np.fft.fft(xr, axis=1)
        assert nl2(d.reshape(count, n), ref) <= 2e-6 * math.log2(n), inv
    f.close()
