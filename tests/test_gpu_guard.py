"""Guard bands: every device entry point writes exactly its output region.
A canary pattern surrounds each image (and fills the gaps between the images
of a batch); after the transforms the pattern must be intact bitwise, and
every input region of an out-of-place call must be unchanged."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GUARD = 4096  # complex elements on each side
CANARY = np.uint32(0x7FA5A5A5)  # a NaN bit pattern no transform produces from finite data


@pytest.fixture(scope="module")
def env():
    import torch

    from math_amd import mdft as M

    M.init(0)
    return M, torch


def guarded(torch, n, rng):
    """Device buffer of n complex values between canary bands; returns (buf, view of the payload)."""
    host = np.empty(2 * (n + 2 * GUARD), np.uint32)
    host[:] = CANARY
    buf = torch.from_numpy(host.view(np.float32).copy()).cuda()
    payload = (rng.uniform(-1, 1, 2 * n)).astype(np.float32)
    buf[2 * GUARD:2 * (GUARD + n)] = torch.from_numpy(payload).cuda()
    return buf


def bands_intact(buf, n):
    a = buf.cpu().numpy().view(np.uint32)
    return bool(np.all(a[:2 * GUARD] == CANARY) and np.all(a[2 * (GUARD + n):] == CANARY))


def ptr(buf, elem_offset=GUARD):
    return buf.data_ptr() + 8 * elem_offset


@pytest.mark.parametrize("w,h", [(1, 1), (1, 37), (53, 1), (4093, 3), (3, 4093), (17, 2049), (640, 480),
                                 (16385, 2), (2, 16385)])
def test_plan_writes_only_its_image(env, w, h):
    M, torch = env
    rng = np.random.default_rng(w * 7 + h)
    n = w * h
    src = guarded(torch, n, rng)
    dst = guarded(torch, n, rng)
    before = src.cpu().numpy().copy()
    p = M.Plan(w, h)
    s = torch.cuda.current_stream().cuda_stream
    p.exec_device(False, ptr(src), ptr(dst), s)   # out of place
    p.exec_device(True, ptr(dst), ptr(dst), s)    # in place
    torch.cuda.synchronize()
    p.close()
    assert bands_intact(dst, n), f"{w}x{h}: write outside the output image"
    assert np.array_equal(src.cpu().numpy().view(np.uint32), before.view(np.uint32)), f"{w}x{h}: input modified"


def test_batch_writes_only_its_images(env):
    M, torch = env
    rng = np.random.default_rng(3)
    sizes = [(37, 61), (1, 9), (512, 3), (3, 700), (100, 100), (1, 1), (2049, 5), (64, 64)]
    invs = [bool(i % 2) for i in range(len(sizes))]
    offs, total = [], GUARD
    for (w, h) in sizes:
        offs.append(total)
        total += w * h + 257  # a canary gap after every image
    host = np.empty(2 * (total + GUARD), np.uint32)
    host[:] = CANARY
    mask = np.zeros(total + GUARD, bool)
    for o, (w, h) in zip(offs, sizes):
        mask[o:o + w * h] = True
        host.view(np.float32)[2 * o:2 * (o + w * h)] = rng.uniform(-1, 1, 2 * w * h).astype(np.float32)
    src = torch.from_numpy(host.view(np.float32).copy()).cuda()
    dst_host = np.empty_like(host)
    dst_host[:] = CANARY
    dst = torch.from_numpy(dst_host.view(np.float32).copy()).cuda()
    for flags in (0, M.MDFT_BATCH_GRAPH | M.MDFT_BATCH_TUNE):
        b = M.Batch(sizes, invs, offs, offs, flags)
        b.exec_device(src.data_ptr(), dst.data_ptr(), torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        b.check()
        b.close()
        out = dst.cpu().numpy().view(np.uint32).reshape(-1, 2)
        assert np.all(out[~mask] == CANARY), f"flags={flags}: write outside the batch's images"
        assert np.array_equal(src.cpu().numpy().view(np.uint32), host), "batch input modified"


def test_fft_and_image_write_only_their_outputs(env):
    M, torch = env
    rng = np.random.default_rng(4)
    count, n = 9, 256
    buf = guarded(torch, count * n, rng)
    f = M.Fft(count, n)
    f.exec_device(False, ptr(buf), ptr(buf), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    f.close()
    assert bands_intact(buf, count * n)
    # image: R32F output of w x h floats inside a canary buffer
    w, h = 123, 45
    pix = torch.rand((h + 3, w + 5, 4), dtype=torch.float32, device="cuda")
    out_host = np.empty(2 * GUARD + w * h, np.uint32)
    out_host[:] = CANARY
    out = torch.from_numpy(out_host.view(np.float32).copy()).cuda()
    img = M.Image(w, h)
    img.exec_device(pix.data_ptr(), w + 5, h + 3, w + 5, 2, 1, out.data_ptr() + 4 * GUARD, w,
                    torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    img.close()
    a = out.cpu().numpy().view(np.uint32)
    assert np.all(a[:GUARD] == CANARY) and np.all(a[GUARD + w * h:] == CANARY)
