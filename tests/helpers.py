"""Shared helpers of the test suite (inputs, tolerances, comparisons)."""
import math

import numpy as np

from oracle import oracle as O


def rand_c64(rng, shape, dist="uniform"):
    if dist == "uniform":
        re = rng.uniform(-1.0, 1.0, shape)
        im = rng.uniform(-1.0, 1.0, shape)
    else:
        re = rng.standard_normal(shape)
        im = rng.standard_normal(shape)
    return (re + 1j * im).astype(np.complex64)


def tol_nl2(w, h):
    """Normalised-L2 bound: 2e-6 * log2(padded N), padded N = compute_m(w) * compute_m(h)."""
    return 2e-6 * max(1.0, math.log2(O.compute_m(w) * O.compute_m(h)))


# the reference's own pass criterion (src/gpu/dft/test/test_dft.cpp:59)
REF_REL_L1 = 1e-4


def check_against_fp64(got, x, inverse, label=""):
    """Compare an FP32 result (h, w) with the FP64 oracle of input x."""
    h, w = x.shape
    ref = O.dft2d_64(x, inverse, nthreads=8)
    m = O.error_metrics(got.reshape(h, w), ref)
    assert m["nl2"] <= tol_nl2(w, h), f"{label} {w}x{h} inv={inverse}: nl2 {m['nl2']:.3e} > {tol_nl2(w, h):.3e}"
    assert m["rel_l1"] <= REF_REL_L1, f"{label} {w}x{h} inv={inverse}: rel L1 {m['rel_l1']:.3e}"
    return m
