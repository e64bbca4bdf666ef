"""CPU tests of the C ABI boundary: the library loads, exports every entry
point include/mdft.h declares, and reports the reference's error messages
without touching a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from math_amd import mdft as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "mdft.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mdft_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("mdft_plan_create", "mdft_exec", "mdft_exec_device", "mdft_plan_destroy", "mdft_last_error",
                 "mdft_batch_create", "mdft_batch_exec_device"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = M.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", M.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\b(mdft_[a-z_0-9]+)\b", out))
    assert set(declared_functions()) <= exported


def test_version():
    assert M.lib().mdft_version().startswith(b"mdft")


def test_size_errors_use_reference_text():
    h = ctypes.c_void_p()
    rc = M.lib().mdft_plan_create(ctypes.byref(h), 0, 5)
    assert rc == M.MDFT_ERR_SIZE
    assert M.lib().mdft_last_error() == b"Error size 0 for compute m"
    rc = M.lib().mdft_plan_create(ctypes.byref(h), 7, (1 << 29) + 1)
    assert rc == M.MDFT_ERR_SIZE
    assert M.lib().mdft_last_error() == b"Error size 536870913 for compute m"
    assert not h.value


def test_exec_without_plan():
    d = np.zeros(6, np.complex64)
    assert M.lib().mdft_exec(None, 0, d.ctypes.data, 6) == M.MDFT_ERR_NO_BUFFERS
    assert M.lib().mdft_last_error() == b"No DFT buffers"
    assert M.lib().mdft_exec_device(None, 0, d.ctypes.data, d.ctypes.data, None) == M.MDFT_ERR_NO_BUFFERS
    with pytest.raises(M.ErrorException, match="No DFT buffers"):
        M.create_compute_vector().exec(False, d)


def test_null_out_pointers():
    assert M.lib().mdft_plan_create(None, 3, 3) == M.MDFT_ERR_INVALID
    assert M.lib().mdft_batch_create(None, None, 0, 0) == M.MDFT_ERR_INVALID
    assert M.lib().mdft_plan_destroy(None) == M.MDFT_OK
    assert M.lib().mdft_batch_destroy(None) == M.MDFT_OK


def test_create_hip_rejects_non_positive():
    with pytest.raises(M.ErrorException):
        M.create_hip(0, 3)


def test_fft_size_errors_use_reference_text():
    # src/gpu/dft/fft.cpp:298-306, checked before any device work
    h = ctypes.c_void_p()
    assert M.lib().mdft_fft_create(ctypes.byref(h), 4, 0) == M.MDFT_ERR_SIZE
    assert M.lib().mdft_last_error() == b"FFT size 0 is not positive"
    assert M.lib().mdft_fft_create(ctypes.byref(h), 4, 12) == M.MDFT_ERR_SIZE
    assert M.lib().mdft_last_error() == b"FFT size 12 is not an integral power of 2"
    assert M.lib().mdft_fft_create(ctypes.byref(h), 4, 1 << 30) == M.MDFT_ERR_SIZE
    assert M.lib().mdft_last_error() == b"FFT size 1073741824 exceeds the longest supported line (2^29)"
    assert M.lib().mdft_fft_create(None, 4, 8) == M.MDFT_ERR_INVALID
    assert M.lib().mdft_fft_destroy(None) == M.MDFT_OK
    d = np.zeros(8, np.complex64)
    assert M.lib().mdft_fft_exec(None, 0, d.ctypes.data, 8) == M.MDFT_ERR_NO_BUFFERS
    with pytest.raises(M.ErrorException, match="not an integral power of 2"):
        M.create_fft(3, 6)


def test_image_create_rejects_bad_format_and_size():
    h = ctypes.c_void_p()
    assert M.lib().mdft_image_create(ctypes.byref(h), 8, 8, 7) == M.MDFT_ERR_INVALID
    assert b"pixel format" in M.lib().mdft_last_error()
    assert M.lib().mdft_image_create(ctypes.byref(h), 0, 8, 0) == M.MDFT_ERR_SIZE
    assert M.lib().mdft_last_error() == b"Error size 0 for compute m"
    assert M.lib().mdft_image_destroy(None) == M.MDFT_OK


def test_tune_load_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("# comment\n3 0 1 1\n99 0 0 1\n")
    with pytest.raises(M.ErrorException, match="bad entry on line 3"):
        M.tune_load(str(p))
    with pytest.raises(M.ErrorException, match="cannot read"):
        M.tune_load(str(tmp_path / "missing.txt"))
