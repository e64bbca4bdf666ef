"""GPU parity of the view's image path (mdft_image_*): luminance of an RGBA
rectangle -> forward 2-D DFT -> |X|/(w*h) into an R32F image
(src/gpu/dft/compute.cpp:91-200, dft_copy_input.comp, dft_copy_output.comp),
against the oracle (FP32 luminance, FP64 DFT).  Bound: normalised L2 of the
magnitudes <= 2e-6*log2(padded N), the bound of the complex transform (the
magnitude map is 1-Lipschitz)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import tol_nl2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mdft():
    from math_amd import mdft as M

    M.init(0)
    return M


def run_image(M, pixels, x, y, w, h, pitch_pad=0, dst_pad=0):
    import torch

    rows, cols = pixels.shape[:2]
    fmt = M.PIXEL_RGBA8 if pixels.dtype == np.uint8 else M.PIXEL_RGBA32F
    src_pitch = cols + pitch_pad
    host = np.zeros((rows, src_pitch, 4), pixels.dtype)
    host[:, :cols] = pixels
    src = torch.from_numpy(host).cuda()
    dst_pitch = w + dst_pad
    dst = torch.full((h, dst_pitch), -7.0, dtype=torch.float32, device="cuda")
    img = M.Image(w, h, fmt)
    img.exec_device(src.data_ptr(), cols, rows, src_pitch, x, y, dst.data_ptr(), dst_pitch,
                    torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    out = dst.cpu().numpy()
    img.close()
    if dst_pad:
        assert np.all(out[:, w:] == -7.0), "wrote outside the rectangle's row"
    return out[:, :w]


def check(got, pixels, x, y, w, h):
    ref = O.image_dft_64(pixels, x, y, w, h, nthreads=8)
    m = O.error_metrics(got.astype(np.complex64), ref.astype(np.complex128))
    assert m["nl2"] <= tol_nl2(w, h), (w, h, m)
    return m


# (192, 70): direct 3 * 2^k rows, 3 * 2^k convolution columns; (1100, 140): 3 * 2^k convolutions on both axes
@pytest.mark.parametrize("w,h", [(640, 480), (256, 256), (97, 61), (1, 37), (53, 1), (1, 1), (2, 3), (1024, 8),
                                 (192, 70), (1100, 140)])
def test_rgba32f_rectangles(mdft, w, h):
    rng = np.random.default_rng(w * 131 + h)
    rows, cols = h + 9, w + 5
    px = rng.uniform(0, 1, (rows, cols, 4)).astype(np.float32)
    got = run_image(mdft, px, 3, 6, w, h, pitch_pad=7, dst_pad=3)
    check(got, px, 3, 6, w, h)


def test_rgba8(mdft):
    rng = np.random.default_rng(88)
    px = rng.integers(0, 256, (300, 400, 4), dtype=np.uint8)
    got = run_image(mdft, px, 10, 20, 320, 240)
    check(got, px, 10, 20, 320, 240)


def test_long_lines_take_the_unfused_path(mdft):
    rng = np.random.default_rng(5)
    px = rng.uniform(0, 1, (3, 20000, 4)).astype(np.float32)
    got = run_image(mdft, px, 0, 0, 20000, 3)
    check(got, px, 0, 0, 20000, 3)


def test_matches_complex_plan(mdft):
    # the fused path runs the same line arithmetic as the complex 2-D plan
    rng = np.random.default_rng(6)
    w, h = 130, 77
    px = rng.uniform(0, 1, (h, w, 4)).astype(np.float32)
    got = run_image(mdft, px, 0, 0, w, h)
    lum = O.image_luminance(px, 0, 0, w, h).astype(np.complex64)
    p = mdft.Plan(w, h)
    p.exec(False, lum)
    mag = np.abs(lum.astype(np.complex128)) / (w * h)
    assert np.abs(got - mag).max() <= 4e-7 * np.abs(mag).max()


def test_rectangle_outside_source_is_rejected(mdft):
    import torch

    src = torch.zeros((10, 10, 4), dtype=torch.float32, device="cuda")
    dst = torch.zeros((8, 8), dtype=torch.float32, device="cuda")
    img = mdft.Image(8, 8)
    with pytest.raises(mdft.ErrorException, match="not inside"):
        img.exec_device(src.data_ptr(), 10, 10, 10, 3, 0, dst.data_ptr(), 8)
    with pytest.raises(mdft.ErrorException, match="pitch"):
        img.exec_device(src.data_ptr(), 10, 10, 9, 0, 0, dst.data_ptr(), 8)
    img.close()
