"""GPU tests of the plan bookkeeping around the transform kernels: padded
lengths reported by mdft_plan_info (compute_m, src/gpu/dft/function.cpp:32-45),
the per-device table cache and its limit, the ordering of long-line plans'
shared scratch across streams, and the server's accounting of resident line
tables against its HBM budget and of inputs produced on a client stream."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import check_against_fp64, rand_c64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    from math_amd import mdft as M

    M.init(0)
    return M, torch


@pytest.mark.parametrize("w,h", [(640, 480), (1024, 1024), (4093, 2047), (1, 17), (2, 1), (6007, 997)])
def test_plan_info_padded_lengths(env, w, h):
    M, _ = env
    p = M.Plan(w, h)
    m1, m2, nbytes = p.info()
    assert (m1, m2) == (O.compute_m(w), O.compute_m(h))
    assert nbytes >= 0
    p.close()


def test_cache_limit_evicts_unused_lengths(env):
    M, torch = env
    lengths = [1003, 1009, 1013, 1019]
    plans = [M.Plan(n, 3) for n in lengths]
    b0, n0, lim0 = M.cache_info()
    assert n0 >= len(lengths) + 1
    keep = plans[0]
    for p in plans[1:]:
        p.close()
    try:
        # limit 0: every length no plan holds is freed; the held ones stay
        M.cache_set_limit(0)
        b1, n1, _ = M.cache_info()
        assert n1 <= n0 - 3  # 1009, 1013 and 1019 are gone; 1003 and 3 are held by `keep`
        assert b1 < b0
        # the held plan still runs correctly
        rng = np.random.default_rng(3)
        x = rand_c64(rng, (3, 1003))
        d = x.ravel().copy()
        keep.exec(False, d)
        check_against_fp64(d, x, False, "after trim")
        # an evicted length is rebuilt on demand
        p = M.Plan(1009, 3)
        d = rand_c64(rng, (3, 1009))
        e = d.ravel().copy()
        p.exec(True, e)
        check_against_fp64(e, d, True, "rebuilt")
        p.close()
    finally:
        keep.close()
        M.cache_set_limit(lim0)


def test_long_line_plan_on_two_streams(env):
    # padded length 65536 (Bluestein): the plan's scratch is shared by both calls
    M, torch = env
    rng = np.random.default_rng(21)
    w, h = 20000, 3
    xa, xb = rand_c64(rng, (h, w)), rand_c64(rng, (h, w))
    a = torch.from_numpy(xa.copy()).cuda()
    b = torch.from_numpy(xb.copy()).cuda()
    ya, yb = torch.empty_like(a), torch.empty_like(b)
    torch.cuda.synchronize()
    p = M.Plan(w, h)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(3):
        p.exec_device(False, a.data_ptr(), ya.data_ptr(), s1.cuda_stream)
        p.exec_device(True, b.data_ptr(), yb.data_ptr(), s2.cuda_stream)
    torch.cuda.synchronize()
    check_against_fp64(ya.cpu().numpy(), xa, False, "stream 1")
    check_against_fp64(yb.cpu().numpy(), xb, True, "stream 2")
    p.close()


def test_server_counts_tables_against_budget(env):
    M, torch = env
    w, h = 1021, 1019  # two Bluestein lengths: tables of ~(5n + 4m) * 8 bytes each
    slot = w * h * 8
    src = torch.zeros(w * h, dtype=torch.complex64, device="cuda")
    torch.cuda.synchronize()
    # a budget that holds the slot alone but not the slot plus the two lengths' tables
    s = M.Server(slot + 4096, streams=1, max_batch=4)
    with pytest.raises(M.ErrorException, match="resident tables"):
        s.submit(w, h, False, src.data_ptr())
    assert s.stats()["table_bytes"] == 0  # a refused request leaves nothing resident
    s.close()
    s = M.Server(4 * slot, streams=1, max_batch=4)
    s.prepare([w, h])
    st = s.stats()
    expect = sum(((5 * n + 4 * O.compute_m(n)) * 8) for n in (w, h))
    assert st["table_bytes"] >= expect
    t = s.submit(w, h, False, src.data_ptr())
    s.wait(t, 30_000_000)
    s.release(t)
    s.close()


def test_server_submit_stream_orders_the_input(env):
    M, torch = env
    rng = np.random.default_rng(8)
    w, h = 700, 500
    x = rand_c64(rng, (h, w))
    host = torch.from_numpy(x.ravel().copy()).pin_memory()
    src = torch.zeros(w * h, dtype=torch.complex64, device="cuda")
    torch.cuda.synchronize()
    s = M.Server(64 << 20, streams=2, max_batch=4)
    producer = torch.cuda.Stream()
    with torch.cuda.stream(producer):
        # a long kernel ahead of the copy keeps the input unwritten for a while
        spin = torch.randn(4096, 4096, device="cuda")
        for _ in range(8):
            spin = spin @ spin
            spin = spin / spin.abs().max()
        src.copy_(host, non_blocking=True)
    t = s.submit_stream(w, h, False, src.data_ptr(), producer.cuda_stream)
    s.wait(t, 60_000_000)
    got = s.fetch(t, w * h).reshape(h, w)
    s.release(t)
    s.close()
    torch.cuda.synchronize()
    check_against_fp64(got, x, False, "submit_stream")
