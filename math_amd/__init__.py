"""mdft: MI355X-native 2-D DFT (drop-in for the 2-D DFT path of cppd/math).

The product is the C ABI library math_amd/lib/libmdft.so (include/mdft.h);
this package holds its sources (csrc/), its build script (build.py) and
ctypes bindings (mdft.py).
"""
from . import mdft  # noqa: F401
