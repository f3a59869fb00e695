"""Thin ctypes binding to the system Zstandard library (libzstd.so.1).

The .asz container is one Zstd frame (level 3) around the Huffman stream.  The
product CLI links the same library from C++; this module serves the Python
harness and tests.
"""
from __future__ import annotations

import ctypes
import ctypes.util

_CANDIDATES = ("/usr/lib/x86_64-linux-gnu/libzstd.so.1", "libzstd.so.1", ctypes.util.find_library("zstd"))
_lib = None
for _c in _CANDIDATES:
    if not _c:
        continue
    try:
        _lib = ctypes.CDLL(_c)
        break
    except OSError:
        continue
if _lib is None:
    raise ImportError("libzstd.so.1 not found")

_lib.ZSTD_compressBound.restype = ctypes.c_size_t
_lib.ZSTD_compressBound.argtypes = [ctypes.c_size_t]
_lib.ZSTD_compress.restype = ctypes.c_size_t
_lib.ZSTD_compress.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
_lib.ZSTD_decompress.restype = ctypes.c_size_t
_lib.ZSTD_decompress.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t]
_lib.ZSTD_getFrameContentSize.restype = ctypes.c_ulonglong
_lib.ZSTD_getFrameContentSize.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
_lib.ZSTD_isError.restype = ctypes.c_uint
_lib.ZSTD_isError.argtypes = [ctypes.c_size_t]


def compress(data: bytes, level: int = 3) -> bytes:
    cap = _lib.ZSTD_compressBound(len(data))
    out = ctypes.create_string_buffer(cap)
    n = _lib.ZSTD_compress(out, cap, data, len(data), level)
    if _lib.ZSTD_isError(n):
        raise RuntimeError("ZSTD_compress failed")
    return out.raw[:n]


def decompress(frame: bytes) -> bytes:
    size = _lib.ZSTD_getFrameContentSize(frame, len(frame))
    if size >= (1 << 63):
        raise RuntimeError("not a Zstd frame with a known content size")
    out = ctypes.create_string_buffer(max(int(size), 1))
    n = _lib.ZSTD_decompress(out, size, frame, len(frame))
    if _lib.ZSTD_isError(n) or n != size:
        raise RuntimeError("ZSTD_decompress error")
    return out.raw[:n]
