"""MI355X-native normal-estimation pipeline (Morton -> sort -> radix tree ->
octree -> exact kNN -> normals) behind a C ABI (include/shapes_c.h).

Importing the package loads ``libshapes_hip.so``; it raises ImportError when
the library has not been built.  There is no CPU fallback.
"""
from ._lib import LIB_PATH, SxError, load

load()

from .api import *  # noqa: E402,F401,F403
from .api import __all__ as _api_all  # noqa: E402
from .device import DeviceArray, Event, Stream, device_count, device_name, fetch, synchronize  # noqa: E402

__all__ = ["LIB_PATH", "SxError", "load", "DeviceArray", "Event", "Stream", "device_count", "device_name", "fetch",
           "synchronize", *_api_all]
