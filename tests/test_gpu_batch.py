"""GPU tests of the persistent batched launch: every transform in a mixed
batch equals the same transform run alone (bitwise), and matches the FP64
oracle within 2e-6*log2(padded N); dependency-counter, two-phase and
hipGraph-replay modes all agree."""
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


def mixed_sizes(rng, count, lo, hi, p_pow2=0.4):
    out = []
    for _ in range(count):
        dims = []
        for _ in range(2):
            if rng.random() < p_pow2:
                dims.append(int(2 ** rng.integers(int(np.ceil(np.log2(lo))), int(np.log2(hi)) + 1)))
            else:
                dims.append(int(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        out.append(tuple(dims))
    return out


def layout(sizes):
    offs, o = [], 0
    for (w, h) in sizes:
        offs.append(o)
        o += w * h
    return offs, o


def run_batch(M, torch, sizes, invs, xs, flags, in_place, replays=1):
    offs, total = layout(sizes)
    flat = np.concatenate([x.ravel() for x in xs]) if xs else np.zeros(0, np.complex64)
    src = torch.from_numpy(flat.copy()).cuda()
    dst = src if in_place else torch.empty_like(src)
    b = M.Batch(sizes, invs, offs, offs, flags)
    s = torch.cuda.current_stream().cuda_stream
    outs = []
    for r in range(replays):
        if in_place and r > 0:
            src.copy_(torch.from_numpy(flat))
        b.exec_device(src.data_ptr(), dst.data_ptr(), s)
        b.check()
        torch.cuda.synchronize()
        outs.append(dst.cpu().numpy().copy())
    b.close()
    return [[o[off:off + w * h].reshape(h, w) for off, (w, h) in zip(offs, sizes)] for o in outs]


def single(M, x, inv):
    h, w = x.shape
    d = x.copy()
    M.Plan(w, h).exec(inv, d)
    return d


@pytest.mark.parametrize("flags,in_place", [(0, False), (0, True), (1, False), (2, False), (3, True), (4, False), (5, True),
                                            (9, False), (8, True)])
def test_batch_matches_single_and_oracle(env, flags, in_place):
    M, torch = env
    rng = np.random.default_rng(100 + flags)
    sizes = mixed_sizes(rng, 48, 17, 700) + [(1, 1), (1, 37), (53, 1), (2, 2)]
    invs = [bool(rng.integers(0, 2)) for _ in sizes]
    xs = [rand_c64(rng, (h, w), "normal") for (w, h) in sizes]
    print(f"batch flags={flags} in_place={in_place} n={len(sizes)}", flush=True)
    results = run_batch(M, torch, sizes, invs, xs, flags, in_place, replays=3)
    for rep in results:
        for (w, h), inv, x, got in zip(sizes, invs, xs, rep):
            alone = single(M, x, inv)
            assert np.array_equal(got, alone), f"{w}x{h} inv={inv}: batch result differs from the single transform"
    for (w, h), inv, x, got in zip(sizes, invs, xs, results[0]):
        check_against_fp64(got, x, inv, "batch")


def test_batch_large_lines(env):
    M, torch = env
    rng = np.random.default_rng(5)
    sizes = [(4097, 3), (3, 4097), (8192, 2), (2, 8192), (2048, 2048), (1500, 900)]
    invs = [False, True, False, True, False, True]
    xs = [rand_c64(rng, (h, w)) for (w, h) in sizes]
    res = run_batch(M, torch, sizes, invs, xs, 0, False)[0]
    for (w, h), inv, x, got in zip(sizes, invs, xs, res):
        assert np.array_equal(got, single(M, x, inv))
        check_against_fp64(got, x, inv, "batch large")


def test_batch_rejects_oversize(env):
    M, _ = env
    with pytest.raises(M.ErrorException, match="batch padded-length limit"):
        M.Batch([(4099, 5)], [False], [0], [0])


def test_batch_direct_three_times_power_of_two(env):
    # lines of exactly 3 * 2^k points (192 .. 6144) run a direct FFT of that length
    M, torch = env
    rng = np.random.default_rng(33)
    sizes = [(192, 768), (768, 192), (1536, 384), (384, 1536), (3072, 6), (6, 3072), (6144, 3), (3, 6144), (96, 192),
             (193, 191)]
    invs = [bool(i % 2) for i in range(len(sizes))]
    xs = [rand_c64(rng, (h, w)) for (w, h) in sizes]
    res = run_batch(M, torch, sizes, invs, xs, 0, False)[0]
    for (w, h), inv, x, got in zip(sizes, invs, xs, res):
        assert np.array_equal(got, single(M, x, inv)), f"{w}x{h}"
        check_against_fp64(got, x, inv, f"batch 3*2^k {w}x{h}")


def test_batch_phases_and_info(env):
    # the two launch phases run separately (mdft_batch_exec_phase) give the
    # same bits as the whole batch; mdft_batch_info accounts every item and
    # the algorithmic bytes 16 * w * h per pass
    M, torch = env
    rng = np.random.default_rng(41)
    sizes = mixed_sizes(rng, 40, 17, 1500) + [(1, 9), (9, 1), (24, 24)]
    invs = [bool(rng.integers(0, 2)) for _ in sizes]
    xs = [rand_c64(rng, (h, w)) for (w, h) in sizes]
    whole = run_batch(M, torch, sizes, invs, xs, 1, False)[0]
    offs, total = layout(sizes)
    src = torch.from_numpy(np.concatenate([x.ravel() for x in xs])).cuda()
    dst = torch.zeros_like(src)
    b = M.Batch(sizes, invs, offs, offs, 1)
    info = b.info()
    s = torch.cuda.current_stream().cuda_stream
    b.exec_phase(1, src.data_ptr(), dst.data_ptr(), s)
    b.exec_phase(2, src.data_ptr(), dst.data_ptr(), s)
    b.check()
    torch.cuda.synchronize()
    got = dst.cpu().numpy()
    b.close()
    for off, (w, h), ref in zip(offs, sizes, whole):
        assert np.array_equal(got[off:off + w * h].reshape(h, w), ref), f"{w}x{h}: phases differ from the batch"
    assert info["items"][0] > 0 and info["items"][1] > 0
    assert info["threads"][0] in (256, 512, 1024) and info["threads"][1] >= info["threads"][0]
    # every image with both passes moves 16*w*h bytes per pass; one-pass images 16*w*h in phase 1
    expect = sum(16 * w * h * ((w > 1) + (h > 1)) for (w, h) in sizes)
    assert sum(info["bytes"]) == expect


@pytest.mark.parametrize("flags", [0, 1])
def test_batch_empty_and_copy_only(env, flags):
    # no transforms, and transforms of 1x1 images only (identities: a copy each)
    M, torch = env
    s = torch.cuda.current_stream().cuda_stream
    buf = torch.zeros(4, dtype=torch.complex64, device="cuda")
    b = M.Batch([], [], [], [], flags)
    b.exec_device(buf.data_ptr(), buf.data_ptr(), s)
    b.check()
    assert b.info()["items"] == (0, 0)
    b.close()
    src = torch.tensor([1 + 2j, 3 - 4j, -5 + 0.5j], dtype=torch.complex64, device="cuda")
    dst = torch.zeros_like(src)
    b = M.Batch([(1, 1)] * 3, [False, True, False], [0, 1, 2], [2, 0, 1], flags)
    b.exec_device(src.data_ptr(), dst.data_ptr(), s)
    b.check()
    torch.cuda.synchronize()
    b.close()
    assert dst.cpu().tolist() == [3 - 4j, -5 + 0.5j, 1 + 2j]


def test_batch_phase_errors(env):
    M, torch = env
    s = torch.cuda.current_stream().cuda_stream
    a = torch.zeros(64 * 64, dtype=torch.complex64, device="cuda")
    c = torch.zeros_like(a)
    b = M.Batch([(64, 64)], [False], [0], [0], M.MDFT_BATCH_PERSISTENT)
    with pytest.raises(M.ErrorException, match="only grid-shape batches run by phase"):
        b.exec_phase(1, a.data_ptr(), a.data_ptr(), s)
    b.close()
    b = M.Batch([(64, 64)], [False], [0], [0], 0)
    with pytest.raises(M.ErrorException, match="phase must be 1 or 2"):
        b.exec_phase(3, a.data_ptr(), a.data_ptr(), s)
    b.exec_phase(1, a.data_ptr(), a.data_ptr(), s)
    with pytest.raises(M.ErrorException, match="needs the pointers of the phase 1 call"):
        b.exec_phase(2, c.data_ptr(), c.data_ptr(), s)
    torch.cuda.synchronize()
    b.close()
