"""multians_amd -- MI355X-native (gfx950) decoder for single-table tANS streams
coded for self-synchronising subsequence decoding (the method of
weissenberger/multians), behind that project's decoder and buffer interfaces.

Native parts (multians_amd/lib/, built by `python -m multians_amd.build`):
  libmultians_amd.so  HIP kernels + C ABI (include/mans_c_api.h)
  libcuhd_amd.so      drop-in C++ classes (include/cuhd_*.h) over the C ABI
"""
from ._lib import MansError, lib  # noqa: F401
from .decoder import DeviceStream, EncodedStream, encode_symbols  # noqa: F401

__all__ = ["MansError", "lib", "DeviceStream", "EncodedStream", "encode_symbols"]
