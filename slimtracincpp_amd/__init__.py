"""MI355X-native per-pixel ray tracer with the application-layer API of SlimTracin.

The product is the C++/HIP library `libslimrt.so` (C ABI in include/slimrt.h, C++ application API in
include/slimtracin/); this package holds its sources (csrc/), the Python host-side driver (runtime.py) and the
seeded synthetic scenes used for tests and measurement (scenes.py, specfile.py).
"""
from . import scenes  # noqa: F401
