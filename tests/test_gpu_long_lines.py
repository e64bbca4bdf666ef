"""The longest lines: padded lengths 2^29 and 2^30 (three-factor steps).

The reference accepts sides up to 2^29 (compute_m, src/gpu/dft/function.cpp:
34-38), which pad to 2^30 for Bluestein.  At these sizes the FP64 oracle is
out of reach, so the device result is checked through properties that do not
depend on the size: sampled bins against an FP64 direct sum on the device
(exact index reduction j*k mod n), the round trip forward then inverse, and
Parseval.  The three-factor steps themselves are checked element by element
against the oracle at smaller sizes in a child process that lowers the split
threshold (tests/long3_check.py).
"""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def mdft():
    from math_amd import mdft as M

    M.init(0)
    return M


def test_three_factor_steps_against_oracle():
    env = dict(os.environ, MDFT_LARGE_SPLIT3_LOGM="15")
    r = subprocess.run([sys.executable, "-u", os.path.join(ROOT, "tests", "long3_check.py")], env=env,
                       capture_output=True, text=True, timeout=110)
    log(r.stdout[-3000:])
    assert r.returncode == 0, r.stderr[-3000:]
    assert r.stdout.strip().endswith("ok")


def fp64_bins(x, ks, inverse):
    """X[k] = sum_j x_j exp(-+2 pi i (j k mod n) / n) in FP64 on the device (1/n for the inverse)."""
    import torch

    n = x.numel()
    out = []
    chunk = 1 << 25
    for k in ks:
        acc = torch.zeros((), dtype=torch.complex128, device=x.device)
        for a in range(0, n, chunk):
            b = min(n, a + chunk)
            j = torch.arange(a, b, dtype=torch.int64, device=x.device)
            ph = ((j * int(k)) % n).to(torch.float64) * ((2.0 if inverse else -2.0) * math.pi / n)
            acc = acc + (x[a:b].to(torch.complex128) * torch.polar(torch.ones_like(ph), ph)).sum()
        out.append(acc / n if inverse else acc)
    return torch.stack(out)


def fp64_bins_2d(lines, ks, inverse):
    """Bins ks along the long axis of every line of a 2-D DFT whose short axis
    holds len(lines) lines: 1-D FP64 bins per line, then the short axis's DFT."""
    import torch

    F = torch.stack([fp64_bins(line, ks, inverse) for line in lines])  # (L, len(ks))
    L = len(lines)
    c = torch.arange(L, dtype=torch.float64, device=F.device)
    W = torch.polar(torch.ones(L, L, dtype=torch.float64, device=F.device),
                    torch.outer(c, c) * ((2.0 if inverse else -2.0) * math.pi / L))
    out = W @ F
    return out / L if inverse else out


# (width, height): a 2^29 power of two (three factors 10/10/9), a Bluestein side
# padding to 2^29, the longest Bluestein side 2^29 - 1 (padding to 2^30), and
# two columns padding to 2^29 (element stride 2 in every step)
@pytest.mark.parametrize("w,h", [(1 << 29, 1), ((1 << 27) + 3, 1), ((1 << 29) - 1, 1), (2, (1 << 27) + 3)])
def test_longest_lines(mdft, w, h):
    import torch

    n, nl = max(w, h), min(w, h)
    padded = O.compute_m(n)
    log(f"longest lines {w}x{h}: padded {padded}")
    g = torch.Generator(device="cuda").manual_seed(n)
    x = torch.complex(torch.rand((h, w), generator=g, device="cuda") * 2 - 1,
                      torch.rand((h, w), generator=g, device="cuda") * 2 - 1)
    p = mdft.Plan(w, h)
    s = torch.cuda.current_stream().cuda_stream
    X = torch.empty_like(x)
    p.exec_device(False, x.data_ptr(), X.data_ptr(), s)
    y = torch.empty_like(x)
    p.exec_device(True, X.data_ptr(), y.data_ptr(), s)
    torch.cuda.synchronize()
    tol = 2e-6 * math.log2(padded * O.compute_m(nl))
    rng = np.random.default_rng(n)
    ks = [0, 1, n - 1, n // 2] + [int(v) for v in rng.integers(0, n, 8)]

    def lines(t):
        return [t[:, c] if h > w else t[c] for c in range(nl)]

    x64 = x.to(torch.complex128)
    norm = torch.linalg.vector_norm(x64).item()  # Parseval: the RMS of the w*h bins
    got = torch.stack([line[ks] for line in lines(X)]).to(torch.complex128)
    err = (got - fp64_bins_2d(lines(x), ks, False)).abs().max().item() / norm
    log(f"  forward: max bin error / RMS {err:.2e} (bound {tol:.2e})")
    assert err <= tol
    back = torch.stack([line[ks[:4]] for line in lines(y)]).to(torch.complex128)
    err_inv = (back - fp64_bins_2d(lines(X), ks[:4], True)).abs().max().item() / (norm / math.sqrt(n * nl))
    log(f"  inverse: max bin error / RMS {err_inv:.2e}")
    assert err_inv <= tol
    rt = (torch.linalg.vector_norm(y.to(torch.complex128) - x64) / norm).item()
    parseval = torch.linalg.vector_norm(X.to(torch.complex128)).item() ** 2 / (n * nl * norm ** 2) - 1.0
    log(f"  round trip {rt:.2e}, Parseval {parseval:.2e}")
    assert rt <= 2 * tol
    assert abs(parseval) <= tol
    p.close()
    del x, X, y, x64
    torch.cuda.empty_cache()


def test_fft1d_longest(mdft):
    # the batched 1-D FFT at its longest length, two lines back to back
    import torch

    count, n = 2, 1 << 29
    log(f"fft1d {count} x {n}")
    g = torch.Generator(device="cuda").manual_seed(11)
    x = torch.complex(torch.rand((count, n), generator=g, device="cuda") * 2 - 1,
                      torch.rand((count, n), generator=g, device="cuda") * 2 - 1)
    f = mdft.Fft(count, n)
    s = torch.cuda.current_stream().cuda_stream
    X = torch.empty_like(x)
    f.exec_device(False, x.data_ptr(), X.data_ptr(), s)
    y = torch.empty_like(x)
    f.exec_device(True, X.data_ptr(), y.data_ptr(), s)
    torch.cuda.synchronize()
    tol = 2e-6 * math.log2(n)
    ks = [0, 1, n - 1, n // 2, 12345, 98765431]
    for c in range(count):
        norm = torch.linalg.vector_norm(x[c].to(torch.complex128)).item()
        err = (X[c][ks].to(torch.complex128) - fp64_bins(x[c], ks, False)).abs().max().item() / norm
        rt = (torch.linalg.vector_norm((y[c] - x[c]).to(torch.complex128)) / norm).item()
        log(f"  line {c}: max bin error / RMS {err:.2e}, round trip {rt:.2e} (bound {tol:.2e})")
        assert err <= tol and rt <= 2 * tol
    f.close()
    del x, X, y
    torch.cuda.empty_cache()
