"""GPU tests of the serving loop: requests submitted while others run are
batched together, and each result is bitwise the same as the transform run
alone and matches the FP64 oracle within 2e-6*log2(padded N); admission by
HBM budget (FIFO, waits for released slots, refuses oversize requests);
trace replay."""
import threading

import numpy as np
import pytest

from tests.helpers import check_against_fp64, rand_c64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    from math_amd import mdft as M

    M.init(0)
    return M, torch


def log_uniform_sizes(rng, count, lo, hi):
    return [(int(np.exp(rng.uniform(np.log(lo), np.log(hi + 1)))), int(np.exp(rng.uniform(np.log(lo), np.log(hi + 1)))))
            for _ in range(count)]


def fetch(server, ticket, w, h):
    """Copy a finished request's result to host memory."""
    return server.fetch(ticket, w * h).reshape(h, w)


def alone(M, torch, x, inv):
    h, w = x.shape
    src = torch.from_numpy(x.ravel().copy()).cuda()
    dst = torch.empty_like(src)
    p = M.Plan(w, h)
    p.exec_device(inv, src.data_ptr(), dst.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return dst.cpu().numpy().reshape(h, w)


def test_concurrent_requests_same_alone_and_in_mix(env):
    M, torch = env
    rng = np.random.default_rng(77)
    sizes = log_uniform_sizes(rng, 60, 32, 600) + [(1, 1), (1, 40), (40, 1), (4096, 2), (2, 4096)]
    invs = [bool(rng.integers(0, 2)) for _ in sizes]
    xs = [rand_c64(rng, (h, w)) for (w, h) in sizes]
    srcs = [torch.from_numpy(x.ravel().copy()).cuda() for x in xs]
    torch.cuda.synchronize()
    s = M.Server(256 << 20, streams=3, max_batch=16)
    tickets = [None] * len(sizes)
    done = threading.Event()

    def producer():
        for i, ((w, h), inv) in enumerate(zip(sizes, invs)):
            tickets[i] = s.submit(w, h, inv, srcs[i].data_ptr())
        done.set()

    th = threading.Thread(target=producer)
    th.start()
    results = []
    for i, (w, h) in enumerate(sizes):
        while tickets[i] is None:
            pass
        ptr, lat = s.wait(tickets[i], 30_000_000)
        assert lat > 0 and ptr
        results.append(fetch(s, tickets[i], w, h))
        s.release(tickets[i])
    th.join()
    st = s.stats()
    assert st["completed"] == len(sizes) and st["in_flight"] == 0 and st["used_bytes"] == 0
    assert st["launches"] >= 1
    s.close()
    for (w, h), inv, x, got in zip(sizes, invs, xs, results):
        assert np.array_equal(got, alone(M, torch, x, inv)), f"{w}x{h}: served result differs from the single run"
        check_against_fp64(got, x, inv, "served")


def test_admission_waits_for_budget(env):
    M, torch = env
    rng = np.random.default_rng(5)
    w, h = 512, 512
    budget = 2 * w * h * 8 + 4096  # room for two requests at a time
    xs = [rand_c64(rng, (h, w)) for _ in range(8)]
    srcs = [torch.from_numpy(x.ravel().copy()).cuda() for x in xs]
    torch.cuda.synchronize()
    s = M.Server(budget, streams=2, max_batch=4)
    tickets = [s.submit(w, h, False, src.data_ptr()) for src in srcs]
    st = s.stats()
    assert st["used_bytes"] <= budget
    for i, t in enumerate(tickets):
        s.wait(t, 30_000_000)
        got = fetch(s, t, w, h)
        s.release(t)
        check_against_fp64(got, xs[i], False, "budgeted")
    assert s.stats()["completed"] == 8
    with pytest.raises(M.ErrorException, match="exceeds the HBM budget"):
        s.submit(1024, 1024, False, srcs[0].data_ptr())
    with pytest.raises(M.ErrorException, match="padded-length limit"):
        s.submit(8193, 2, False, srcs[0].data_ptr())
    s.close()


def test_replay_trace(env):
    M, torch = env
    rng = np.random.default_rng(9)
    n = 200
    sizes = log_uniform_sizes(rng, n, 32, 512)
    pool = torch.randn(512 * 512, dtype=torch.complex64, device="cuda")
    torch.cuda.synchronize()
    arrivals = np.cumsum(rng.exponential(20.0, n))
    s = M.Server(1 << 30)
    s.prepare(sorted({d for wh in sizes for d in wh}))
    lat, makespan = s.replay([w for w, _ in sizes], [h for _, h in sizes], [i % 2 == 0 for i in range(n)],
                             [pool.data_ptr()] * n, arrivals)
    assert lat.shape == (n,) and np.all(lat > 0) and makespan >= arrivals[-1]
    st = s.stats()
    assert st["completed"] == n and st["in_flight"] == 0
    s.close()


def test_batching_window(env):
    # with a window (hold requests while another micro-batch runs until 8 wait
    # or the oldest has waited 300 us) every result is still the alone result
    M, torch = env
    rng = np.random.default_rng(91)
    sizes = log_uniform_sizes(rng, 80, 32, 400)
    invs = [bool(rng.integers(0, 2)) for _ in sizes]
    xs = [rand_c64(rng, (h, w)) for (w, h) in sizes]
    srcs = [torch.from_numpy(x.ravel().copy()).cuda() for x in xs]
    torch.cuda.synchronize()
    s = M.Server(256 << 20, streams=2, max_batch=32)
    s.set_batching(8, 300.0)
    tickets = [s.submit(w, h, inv, src.data_ptr()) for (w, h), inv, src in zip(sizes, invs, srcs)]
    for t, (w, h), inv, x in zip(tickets, sizes, invs, xs):
        s.wait(t, 30_000_000)
        got = fetch(s, t, w, h)
        s.release(t)
        assert np.array_equal(got, alone(M, torch, x, inv)), f"{w}x{h}: windowed result differs from alone"
    st = s.stats()
    s.close()
    assert st["completed"] == len(sizes)
    assert st["launches"] < len(sizes)
    with pytest.raises(M.ErrorException):
        M.Server(1 << 20, streams=1, max_batch=4).set_batching(2, -1.0)
