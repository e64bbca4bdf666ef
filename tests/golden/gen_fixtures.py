"""Writes tests/golden/dft2d_fixtures.npz: small 2-D and 1-D DFT vectors.

Inputs are complex64 (seeded numpy); expected outputs are numpy's FP64 FFT
of the upcast inputs with the reference's semantics: forward unnormalised,
inverse scaled by 1/(w*h) (src/dft/fftw.cpp:140-150).  numpy's pocketfft is
an independent FP64 implementation, so these vectors pin the oracle.  Also
the constant-data case of src/gpu/dft/test/test_dft.cpp:391-400.

Run: python tests/golden/gen_fixtures.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

SIZES_2D = [(2, 3), (1, 1), (1, 7), (7, 1), (3, 5), (5, 3), (4, 4), (6, 10), (13, 17), (17, 13), (16, 32), (31, 29),
            (37, 61), (100, 3), (3, 100), (24, 20), (64, 2)]
LINES_1D = [1, 2, 3, 5, 17, 127, 128, 640, 997]


def main():
    rng = np.random.default_rng(20261202)
    arrays = {}
    cases = []
    for (w, h) in SIZES_2D:
        cases.append((w, h))
    for n in LINES_1D:
        cases.append((n, 1))
    for i, (w, h) in enumerate(cases):
        if (w, h) == (2, 3) and i == 0:
            x = np.array([1, 2, 3, 4, 5, 6], np.complex64).reshape(3, 2)
        else:
            x = (rng.uniform(-1, 1, (h, w)) + 1j * rng.uniform(-1, 1, (h, w))).astype(np.complex64)
        x64 = x.astype(np.complex128)
        arrays[f"in_{i}"] = x
        arrays[f"fwd_{i}"] = np.fft.fft2(x64)
        arrays[f"inv_{i}"] = np.fft.ifft2(x64)
    arrays["sizes"] = np.array(cases, np.int64)
    path = os.path.join(HERE, "dft2d_fixtures.npz")
    np.savez_compressed(path, **arrays)
    print(f"wrote {path}: {len(cases)} cases")


if __name__ == "__main__":
    main()
