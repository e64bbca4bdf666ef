"""Callers may record the device entry points into their own HIP graph (the
analogue of recording Dft::compute_commands into a command buffer,
src/gpu/dft/dft.h:41): a transform, a long-line transform (plan scratch) and a
batch captured into a torch CUDA graph give, on every replay with eager work
in between, bitwise the result of the eager call."""
import numpy as np
import pytest

from tests.helpers import rand_c64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    from math_amd import mdft as M

    M.init(0)
    return M, torch


@pytest.mark.parametrize("w,h", [(640, 480), (1024, 1024), (17, 2049), (20000, 3)])
def test_captured_plan_replays_match_eager(env, w, h):
    M, torch = env
    rng = np.random.default_rng(w + h)
    x = torch.from_numpy(rand_c64(rng, (h, w)).ravel()).cuda()
    y = torch.empty_like(x)
    p = M.Plan(w, h)
    s = torch.cuda.current_stream().cuda_stream
    p.exec_device(False, x.data_ptr(), y.data_ptr(), s)
    torch.cuda.synchronize()
    eager = y.cpu().numpy().copy()

    g = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(cap):
        y.zero_()
        with torch.cuda.graph(g, stream=cap):
            p.exec_device(False, x.data_ptr(), y.data_ptr(), cap.cuda_stream)
    torch.cuda.synchronize()
    for rep in range(3):
        y.fill_(float(rep + 7))  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(y.cpu().numpy(), eager), f"{w}x{h}: replay {rep} differs from the eager call"
    p.close()


def test_captured_batch_replays_match_eager(env):
    M, torch = env
    rng = np.random.default_rng(99)
    sizes = [(37, 61), (300, 7), (1, 20), (512, 512), (100, 1), (257, 129)]
    invs = [False, True, False, True, True, False]
    offs, tot = [], 0
    for (w, h) in sizes:
        offs.append(tot)
        tot += w * h
    x = torch.from_numpy(rand_c64(rng, (tot,))).cuda()
    y = torch.empty_like(x)
    b = M.Batch(sizes, invs, offs, offs, 0)
    b.exec_device(x.data_ptr(), y.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    eager = y.cpu().numpy().copy()
    g = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(cap):
        with torch.cuda.graph(g, stream=cap):
            b.exec_device(x.data_ptr(), y.data_ptr(), cap.cuda_stream)
    torch.cuda.synchronize()
    for rep in range(3):
        y.fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(y.cpu().numpy(), eager), f"batch replay {rep} differs from the eager call"
    b.close()
