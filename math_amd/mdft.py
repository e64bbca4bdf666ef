"""Python bindings (ctypes) over the C ABI of libmdft.so (include/mdft.h).

Mirrors the reference's interfaces for the 2-D DFT path:
  * ComputeVector.create_buffers(width, height) / exec(inverse, data)
    (src/gpu/dft/compute.h:52-60, compute.cpp:229-283),
  * create_hip(x, y).exec(inverse, data), the ns::dft::DFT shape
    (src/dft/dft.h:25-31, src/dft/fftw.cpp:178-186),
plus device-side entry points (Plan.exec_device, Batch) that take raw device
pointers and a HIP stream handle (e.g. a torch tensor's data_ptr() and
torch.cuda.current_stream().cuda_stream).

The library is required: importing this module raises if libmdft.so is not
built, and there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmdft.so")

MDFT_OK = 0
MDFT_ERR_INVALID = -1
MDFT_ERR_SIZE = -2
MDFT_ERR_NO_BUFFERS = -3
MDFT_ERR_DATA_SIZE = -4
MDFT_ERR_DEVICE = -5
MDFT_ERR_BUDGET = -6
MDFT_ERR_TIMEOUT = -7
MDFT_ERR_STATE = -8

MDFT_BATCH_GRAPH = 1
MDFT_BATCH_PERSISTENT = 2
MDFT_BATCH_DEPENDENT = 4
MDFT_BATCH_TUNE = 8


class ErrorException(RuntimeError):
    """Raised with the library's message text (the reference's ns::ErrorException)."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


class mdft_desc(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("inverse", ctypes.c_int32),
        ("reserved", ctypes.c_uint32),
        ("src_offset", ctypes.c_uint64),
        ("dst_offset", ctypes.c_uint64),
    ]


class mdft_server_counters(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("completed", "pending", "in_flight", "launches", "launched_requests",
                                                "budget_bytes", "used_bytes", "table_bytes")]


class mdft_request(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("inverse", ctypes.c_int32),
        ("reserved", ctypes.c_uint32),
        ("src", ctypes.c_void_p),
        ("arrival_us", ctypes.c_double),
    ]


_vp = ctypes.c_void_p
_L = None


def _share_runtime_with_torch():
    # libmdft.so needs libamdhip64.so.7.  torch's HIP library asks for its own
    # bundled copy by the unversioned name, so if libmdft.so loaded first the
    # process would end up with two HIP runtimes and torch would see no GPU.
    # Loading torch first makes libmdft.so bind to the runtime already loaded.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    global _L
    if _L is None:
        _share_runtime_with_torch()
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libmdft.so is not built ({LIB_PATH}); run `python -m math_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "mdft_last_error": ([], ctypes.c_char_p),
            "mdft_version": ([], ctypes.c_char_p),
            "mdft_init": ([ctypes.c_int], ctypes.c_int),
            "mdft_plan_create": ([ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.c_uint32], ctypes.c_int),
            "mdft_plan_destroy": ([_vp], ctypes.c_int),
            "mdft_plan_info": ([_vp, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
            "mdft_exec": ([_vp, ctypes.c_int, _vp, ctypes.c_size_t], ctypes.c_int),
            "mdft_exec_device": ([_vp, ctypes.c_int, _vp, _vp, _vp], ctypes.c_int),
            "mdft_exec_device_noalloc": ([ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _vp, _vp, _vp],
                                         ctypes.c_int),
            "mdft_cache_info": ([ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint32),
                                 ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
            "mdft_cache_set_limit": ([ctypes.c_uint64], ctypes.c_int),
            "mdft_tune_get": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
            "mdft_tune_set": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
            "mdft_tune_auto": ([ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
            "mdft_tune_save": ([ctypes.c_char_p], ctypes.c_int),
            "mdft_tune_load": ([ctypes.c_char_p], ctypes.c_int),
            "mdft_fft_create": ([ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.c_uint32], ctypes.c_int),
            "mdft_fft_exec": ([_vp, ctypes.c_int, _vp, ctypes.c_size_t], ctypes.c_int),
            "mdft_fft_exec_device": ([_vp, ctypes.c_int, _vp, _vp, _vp], ctypes.c_int),
            "mdft_fft_destroy": ([_vp], ctypes.c_int),
            "mdft_image_create": ([ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32],
                                  ctypes.c_int),
            "mdft_image_exec_device": ([_vp, _vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32,
                                        ctypes.c_uint32, _vp, ctypes.c_uint64, _vp], ctypes.c_int),
            "mdft_image_destroy": ([_vp], ctypes.c_int),
            "mdft_batch_create": ([ctypes.POINTER(_vp), ctypes.POINTER(mdft_desc), ctypes.c_uint32, ctypes.c_uint32],
                                  ctypes.c_int),
            "mdft_batch_exec_device": ([_vp, _vp, _vp, _vp], ctypes.c_int),
            "mdft_batch_check": ([_vp], ctypes.c_int),
            "mdft_batch_exec_phase": ([_vp, ctypes.c_int, _vp, _vp, _vp], ctypes.c_int),
            "mdft_batch_info": ([_vp, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                 ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                 ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)], ctypes.c_int),
            "mdft_batch_destroy": ([_vp], ctypes.c_int),
            "mdft_server_create": ([ctypes.POINTER(_vp), ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32],
                                   ctypes.c_int),
            "mdft_server_set_batching": ([_vp, ctypes.c_uint32, ctypes.c_double], ctypes.c_int),
            "mdft_server_prepare": ([_vp, ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32], ctypes.c_int),
            "mdft_server_submit": ([_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _vp,
                                    ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
            "mdft_server_submit_stream": ([_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _vp, _vp,
                                           ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
            "mdft_server_wait": ([_vp, ctypes.c_uint64, ctypes.c_int64, ctypes.POINTER(_vp),
                                  ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
            "mdft_server_release": ([_vp, ctypes.c_uint64], ctypes.c_int),
            "mdft_server_fetch": ([_vp, ctypes.c_uint64, _vp, ctypes.c_size_t], ctypes.c_int),
            "mdft_server_stats": ([_vp, ctypes.POINTER(mdft_server_counters)], ctypes.c_int),
            "mdft_server_replay": ([_vp, ctypes.POINTER(mdft_request), ctypes.c_uint32,
                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
            "mdft_server_destroy": ([_vp], ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _L = L
    return _L


def _check(rc):
    if rc != MDFT_OK:
        msg = lib().mdft_last_error().decode()
        raise ErrorException(rc, msg)


def init(device=0):
    _check(lib().mdft_init(device))


def exec_device_noalloc(width, height, inverse, src_ptr, dst_ptr, stream=0):
    """Allocation-free transform of a device image (mdft_exec_device_noalloc);
    the tables of both sides must be resident (a live plan of the size)."""
    _check(lib().mdft_exec_device_noalloc(int(width), int(height), int(bool(inverse)), src_ptr, dst_ptr, stream))


def cache_info():
    """(bytes, lengths, limit) of this thread's device table cache (mdft_cache_info)."""
    b, n, lim = ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_uint64()
    _check(lib().mdft_cache_info(ctypes.byref(b), ctypes.byref(n), ctypes.byref(lim)))
    return b.value, n.value, lim.value


def cache_set_limit(nbytes):
    _check(lib().mdft_cache_set_limit(int(nbytes)))


def tune_get(shape, blue, cols):
    v = ctypes.c_int()
    _check(lib().mdft_tune_get(shape, int(blue), int(cols), ctypes.byref(v)))
    return v.value


def tune_set(shape, blue, cols, variant):
    _check(lib().mdft_tune_set(shape, int(blue), int(cols), int(variant)))


def tune_auto():
    """Time both tile shapes of every single-image kernel and keep the faster
    (mdft_tune_auto); returns the seconds taken."""
    sec = ctypes.c_double()
    _check(lib().mdft_tune_auto(ctypes.byref(sec)))
    return sec.value


def tune_save(path):
    _check(lib().mdft_tune_save(os.fsencode(path)))


def tune_load(path):
    _check(lib().mdft_tune_load(os.fsencode(path)))


class Plan:
    """One width x height transform (mdft_plan_create)."""

    def __init__(self, width, height):
        self._h = _vp()
        _check(lib().mdft_plan_create(ctypes.byref(self._h), int(width), int(height)))
        self.width, self.height = int(width), int(height)

    def close(self):
        if self._h:
            lib().mdft_plan_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        m1, m2, b = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        _check(lib().mdft_plan_info(self._h, ctypes.byref(m1), ctypes.byref(m2), ctypes.byref(b)))
        return m1.value, m2.value, b.value

    def exec(self, inverse, data):
        """In place on a C-contiguous complex64 numpy array of width*height values."""
        if not isinstance(data, np.ndarray) or data.dtype != np.complex64 or not data.flags.c_contiguous:
            raise ErrorException(MDFT_ERR_DATA_SIZE, "Wrong DFT buffer size")
        _check(lib().mdft_exec(self._h, int(bool(inverse)), data.ctypes.data, data.size))

    def exec_device(self, inverse, src_ptr, dst_ptr, stream=0):
        _check(lib().mdft_exec_device(self._h, int(bool(inverse)), src_ptr, dst_ptr, stream))


class ComputeVector:
    """Shape of ns::gpu::dft::ComputeVector (src/gpu/dft/compute.h:52-60)."""

    def __init__(self):
        self._plan = None

    def create_buffers(self, width, height):
        if self._plan is not None:
            self._plan.close()
            self._plan = None
        self._plan = Plan(width, height)

    def exec(self, inverse, data):
        if self._plan is None:
            _check_null_plan_exec(inverse, data)
        self._plan.exec(inverse, data)


def _check_null_plan_exec(inverse, data):
    # the library raises "No DFT buffers" for a NULL plan (compute.cpp:259-262)
    p = data.ctypes.data if isinstance(data, np.ndarray) else None
    n = data.size if isinstance(data, np.ndarray) else 0
    _check(lib().mdft_exec(None, int(bool(inverse)), p, n))


def create_compute_vector():
    """create_compute_vector(DeviceSearchType) (src/gpu/dft/compute.h:68)."""
    return ComputeVector()


class DFT:
    """Shape of ns::dft::DFT with a fixed size (src/dft/dft.h:25-31)."""

    def __init__(self, x, y):
        if not (x > 0 and y > 0):
            raise ErrorException(MDFT_ERR_SIZE, "Error HIP data size")
        self._plan = Plan(x, y)

    def exec(self, inverse, data):
        self._plan.exec(inverse, data)


def create_hip(x, y):
    """The HIP counterpart of create_fftw(x, y) / create_cufft(x, y) (src/dft/fftw.h:28)."""
    return DFT(x, y)


class Fft:
    """count power-of-two FFTs of length n back to back (mdft_fft_create); the
    shape of ns::gpu::dft::Fft (src/gpu/dft/fft.h:32-53)."""

    def __init__(self, count, n):
        self._h = _vp()
        _check(lib().mdft_fft_create(ctypes.byref(self._h), int(count), int(n)))
        self.count, self.n = int(count), int(n)

    def run_for_data(self, inverse, data):
        """In place on a C-contiguous complex64 numpy array of count*n values."""
        if not isinstance(data, np.ndarray) or data.dtype != np.complex64 or not data.flags.c_contiguous:
            raise ErrorException(MDFT_ERR_DATA_SIZE, "FFT buffer must be a contiguous complex64 array")
        _check(lib().mdft_fft_exec(self._h, int(bool(inverse)), data.ctypes.data, data.size))

    def exec_device(self, inverse, src_ptr, dst_ptr, stream=0):
        _check(lib().mdft_fft_exec_device(self._h, int(bool(inverse)), src_ptr, dst_ptr, stream))

    def close(self):
        if self._h:
            lib().mdft_fft_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def create_fft(count, n):
    """create_fft(device, family_indices, count, n) (src/gpu/dft/fft.h:49-53)."""
    return Fft(count, n)


PIXEL_RGBA32F = 0
PIXEL_RGBA8 = 1


class Image:
    """The view's image path (mdft_image_*): luminance of an RGBA rectangle,
    forward 2-D DFT, |X|/(w*h) into an R32F image; the shape of
    ns::gpu::dft::ComputeImage (src/gpu/dft/compute.h:35-50)."""

    def __init__(self, width, height, pixel_format=PIXEL_RGBA32F):
        self._h = _vp()
        _check(lib().mdft_image_create(ctypes.byref(self._h), int(width), int(height), int(pixel_format)))
        self.width, self.height = int(width), int(height)

    def exec_device(self, src_ptr, src_width, src_height, src_pitch, x, y, dst_ptr, dst_pitch, stream=0):
        _check(lib().mdft_image_exec_device(self._h, src_ptr, int(src_width), int(src_height), int(src_pitch), int(x),
                                            int(y), dst_ptr, int(dst_pitch), stream))

    def close(self):
        if self._h:
            lib().mdft_image_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many transforms of mixed sizes in one persistent launch (mdft_batch_create)."""

    def __init__(self, sizes, inverse, src_offsets, dst_offsets, flags=0):
        n = len(sizes)
        arr = (mdft_desc * max(n, 1))()
        for i, ((w, h), inv, so, do) in enumerate(zip(sizes, inverse, src_offsets, dst_offsets)):
            arr[i] = mdft_desc(int(w), int(h), int(bool(inv)), 0, int(so), int(do))
        self._h = _vp()
        _check(lib().mdft_batch_create(ctypes.byref(self._h), arr, n, flags))

    def exec_device(self, src_ptr, dst_ptr, stream=0):
        _check(lib().mdft_batch_exec_device(self._h, src_ptr, dst_ptr, stream))

    def exec_phase(self, phase, src_ptr, dst_ptr, stream=0):
        """Enqueue launch phase 1 (whole images, row tiles) or 2 (column tiles)."""
        _check(lib().mdft_batch_exec_phase(self._h, int(phase), src_ptr, dst_ptr, stream))

    def info(self):
        """dict: items, algorithmic bytes and tile threads of each phase (mdft_batch_info)."""
        i1, i2, b1, b2, t1, t2 = (ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint64(),
                                  ctypes.c_uint32(), ctypes.c_uint32())
        _check(lib().mdft_batch_info(self._h, ctypes.byref(i1), ctypes.byref(i2), ctypes.byref(b1), ctypes.byref(b2),
                                     ctypes.byref(t1), ctypes.byref(t2)))
        return {"items": (i1.value, i2.value), "bytes": (b1.value, b2.value), "threads": (t1.value, t2.value)}

    def check(self):
        _check(lib().mdft_batch_check(self._h))

    def close(self):
        if self._h:
            lib().mdft_batch_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching server with HBM-budget admission (mdft_server_*)."""

    def __init__(self, budget_bytes, streams=0, max_batch=0):
        self._h = _vp()
        _check(lib().mdft_server_create(ctypes.byref(self._h), int(budget_bytes), int(streams), int(max_batch)))

    def set_batching(self, min_batch, max_wait_us):
        """Batching window of micro-batches launched while another runs (mdft_server_set_batching)."""
        _check(lib().mdft_server_set_batching(self._h, int(min_batch), float(max_wait_us)))

    def prepare(self, lengths):
        arr = (ctypes.c_uint32 * max(len(lengths), 1))(*[int(n) for n in lengths])
        _check(lib().mdft_server_prepare(self._h, arr, len(lengths)))

    def submit(self, width, height, inverse, src_ptr):
        t = ctypes.c_uint64()
        _check(lib().mdft_server_submit(self._h, int(width), int(height), int(bool(inverse)), src_ptr,
                                        ctypes.byref(t)))
        return t.value

    def submit_stream(self, width, height, inverse, src_ptr, stream):
        """Submit an input still being produced on `stream` (mdft_server_submit_stream)."""
        t = ctypes.c_uint64()
        _check(lib().mdft_server_submit_stream(self._h, int(width), int(height), int(bool(inverse)), src_ptr,
                                               stream, ctypes.byref(t)))
        return t.value

    def wait(self, ticket, timeout_us=-1):
        """Returns (device address of the result, latency in microseconds)."""
        r, lat = _vp(), ctypes.c_double()
        _check(lib().mdft_server_wait(self._h, ticket, int(timeout_us), ctypes.byref(r), ctypes.byref(lat)))
        return r.value, lat.value

    def fetch(self, ticket, count):
        """Copy a finished result to a new complex64 host array of `count` values."""
        out = np.empty(count, np.complex64)
        _check(lib().mdft_server_fetch(self._h, ticket, out.ctypes.data, count))
        return out

    def release(self, ticket):
        _check(lib().mdft_server_release(self._h, ticket))

    def stats(self):
        st = mdft_server_counters()
        _check(lib().mdft_server_stats(self._h, ctypes.byref(st)))
        return {n: getattr(st, n) for n, _ in mdft_server_counters._fields_}

    def replay(self, widths, heights, inverses, src_ptrs, arrivals_us):
        n = len(widths)
        reqs = (mdft_request * max(n, 1))()
        for i in range(n):
            reqs[i] = mdft_request(int(widths[i]), int(heights[i]), int(bool(inverses[i])), 0, int(src_ptrs[i]),
                                   float(arrivals_us[i]))
        lat = (ctypes.c_double * max(n, 1))()
        mk = ctypes.c_double()
        _check(lib().mdft_server_replay(self._h, reqs, n, lat, ctypes.byref(mk)))
        return np.frombuffer(lat, dtype=np.float64, count=n).copy(), mk.value

    def close(self):
        if self._h:
            lib().mdft_server_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
