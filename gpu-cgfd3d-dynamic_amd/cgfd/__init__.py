"""Python side of the MI355X-native time-stepping core.

``cgfd.lib``      ctypes bindings of the C ABI (include/cgfd_core.h, cgfd_setup.h)
``cgfd.model``    scenario -> host setup -> descriptor; GPU core wrapper
``cgfd.configs``  seeded synthetic configurations (grids, media, faults)
``cgfd.build``    build recipes for the native libraries
"""
from .lib import load_library, CgfdError  # noqa: F401
