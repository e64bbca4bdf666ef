"""ctypes loader for the CPU oracle (oracle/liboracle.so) and, when built,
the reference's own plan math (oracle/_ref/libref_function.so).

TEST INFRASTRUCTURE ONLY: imported by tests, the smoke check and the
benchmark's CPU-baseline leg as the checker; the product never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(path)
        L.orc_compute_m.argtypes = [ctypes.c_int]
        L.orc_compute_m.restype = ctypes.c_int
        L.orc_compute_h.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp]
        L.orc_compute_h2.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_shared_size.argtypes = [ctypes.c_uint, ctypes.c_uint]
        L.orc_shared_size.restype = ctypes.c_uint
        L.orc_fft32.argtypes = [ctypes.c_int, ctypes.c_int, _fp, ctypes.c_int, ctypes.c_uint]
        L.orc_mul_data32.argtypes = [_fp, ctypes.c_longlong, ctypes.c_int, ctypes.c_int]
        L.orc_diagonals32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint, _fp, _fp, _fp, _fp]
        L.orc_dft2d_ref32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp, ctypes.c_uint, ctypes.c_int]
        L.orc_dft2d_ref32.restype = ctypes.c_int
        L.orc_dft1d_64.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_dft1d_direct64.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_dft2d_64.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.orc_dft2d_64.restype = ctypes.c_int
        L.orc_dft2d_64_from32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp, _dp, ctypes.c_int]
        L.orc_dft2d_64_from32.restype = ctypes.c_int
        L.orc_error_metrics.argtypes = [ctypes.c_size_t, _fp, _dp, _dp]
        L.orc_fill.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t, _fp]
        L.orc_image_luminance.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, _fp]
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's function.cpp compiled as-is, or None when not built."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libref_function.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        R.ref_compute_m.argtypes = [ctypes.c_int]
        R.ref_compute_m.restype = ctypes.c_int
        R.ref_compute_h.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp]
        R.ref_compute_h2.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        R.ref_shared_size.argtypes = [ctypes.c_uint, ctypes.c_uint]
        R.ref_shared_size.restype = ctypes.c_uint
        R.ref_group_size.argtypes = [ctypes.c_uint] * 4
        R.ref_group_size.restype = ctypes.c_uint
        _REF = R
    return _REF


def _c128(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a, a.ctypes.data_as(_dp)


def _c64(a):
    a = np.ascontiguousarray(a, dtype=np.complex64)
    return a, a.ctypes.data_as(_fp)


def compute_m(n):
    return lib().orc_compute_m(int(n))


def compute_h(n, inverse, coef):
    out = np.empty(n, np.complex128)
    lib().orc_compute_h(n, int(bool(inverse)), float(coef), out.ctypes.data_as(_dp))
    return out


def compute_h2(n, m, h):
    h, hp = _c128(h)
    out = np.empty(m, np.complex128)
    lib().orc_compute_h2(n, m, hp, out.ctypes.data_as(_dp))
    return out


def fft32(data, n, inverse, shared_size):
    """Reference FP32 batched radix-2 FFT on a copy of data (count*n complex64)."""
    d, p = _c64(np.array(data, dtype=np.complex64, copy=True))
    lib().orc_fft32(d.size // n, n, p, int(bool(inverse)), shared_size)
    return d


def diagonals32(n1, n2, lds_bytes=65536):
    m1, m2 = compute_m(n1), compute_m(n2)
    t = [np.empty(m1, np.complex64), np.empty(m1, np.complex64), np.empty(m2, np.complex64), np.empty(m2, np.complex64)]
    lib().orc_diagonals32(n1, n2, lds_bytes, *[a.ctypes.data_as(_fp) for a in t])
    return t


def dft2d_ref32(x, inverse, lds_bytes=65536, nthreads=1):
    """FP32 restatement of the reference GPU pipeline. x: (h, w) complex64."""
    h, w = x.shape
    d, p = _c64(np.array(x, dtype=np.complex64, copy=True))
    rc = lib().orc_dft2d_ref32(w, h, int(bool(inverse)), p, lds_bytes, nthreads)
    if rc != 0:
        raise ValueError("Error size for compute m")
    return d.reshape(h, w)


def dft2d_64(x, inverse, nthreads=1):
    """FP64 2-D DFT with the reference's semantics. x: (h, w) complex."""
    x = np.asarray(x)
    h, w = x.shape
    out = np.empty((h, w), np.complex128)
    if x.dtype == np.complex64:
        xx = np.ascontiguousarray(x)
        rc = lib().orc_dft2d_64_from32(w, h, int(bool(inverse)), xx.ctypes.data_as(_fp), out.ctypes.data_as(_dp), nthreads)
    else:
        xx, p = _c128(x)
        rc = lib().orc_dft2d_64(w, h, int(bool(inverse)), p, out.ctypes.data_as(_dp), nthreads)
    if rc != 0:
        raise ValueError("bad size")
    return out


def dft1d_64(x, inverse):
    x, p = _c128(x)
    out = np.empty_like(x)
    lib().orc_dft1d_64(x.size, int(bool(inverse)), p, out.ctypes.data_as(_dp))
    return out


def dft1d_direct64(x, inverse):
    x, p = _c128(x)
    out = np.empty_like(x)
    lib().orc_dft1d_direct64(x.size, int(bool(inverse)), p, out.ctypes.data_as(_dp))
    return out


def error_metrics(got, ref):
    """(normalised L2, normalised max, max abs, reference relative L1)."""
    g, gp = _c64(got)
    r, rp = _c128(ref)
    out = np.zeros(4, np.float64)
    lib().orc_error_metrics(g.size, gp, rp, out.ctypes.data_as(_dp))
    return {"nl2": out[0], "nmax": out[1], "max_abs": out[2], "rel_l1": out[3]}


def fill(seed, dist, shape):
    """Counter-based synthetic data: dist 'uniform' -> U[-1,1), 'normal' -> N(0,1)."""
    out = np.empty(shape, np.complex64)
    lib().orc_fill(seed, 0 if dist == "uniform" else 1, out.size, out.ctypes.data_as(_fp))
    return out


def image_luminance(pixels, x, y, w, h):
    """Luminance of the w x h rectangle at (x, y) of an RGBA image given as a
    (rows, cols, 4) float32 or uint8 array (orc_image_luminance)."""
    pixels = np.ascontiguousarray(pixels)
    fmt = 1 if pixels.dtype == np.uint8 else 0
    if fmt == 0 and pixels.dtype != np.float32:
        raise ValueError("pixels must be float32 or uint8")
    out = np.empty((h, w), np.float32)
    lib().orc_image_luminance(fmt, pixels.ctypes.data, pixels.shape[1], x, y, w, h, out.ctypes.data_as(_fp))
    return out


def image_dft_64(pixels, x, y, w, h, nthreads=1):
    """The view's image path in FP64 after the FP32 luminance: |DFT(lum)| / (w*h)
    (src/gpu/dft/compute.cpp:146-176, dft_copy_output.comp:34-48)."""
    lum = image_luminance(pixels, x, y, w, h).astype(np.complex128)
    return np.abs(dft2d_64(lum, False, nthreads=nthreads)) / (float(w) * float(h))


def tolerance(w, h):
    """Normalised-L2 bound 2e-6 * log2(padded N), padded N = compute_m(w)*compute_m(h)."""
    import math
    pn = compute_m(w) * compute_m(h)
    return 2e-6 * max(1.0, math.log2(pn))
