"""Threaded Zstd frame (asz_zstd_compress_mt) against the single-threaded
level-3 call of compress.tpp:294-299, through the C ABI (host only).

Huffman output is stored raw by Zstd, so the spliced frame of incompressible
input must equal the single-threaded frame byte for byte; compressible input
(a later slice with compressed blocks) must fall back to one call; every frame
must carry its content size (decompress.tpp:83-86 reads it) and decode to the
input.
"""
import ctypes

import numpy as np
import pytest

from asz_amd import _lib, codec

MiB = 1 << 20


def _mt(data: bytes, level=3, threads=4):
    L = _lib.lib()
    buf = np.frombuffer(data, np.uint8)
    out, n, used = ctypes.c_void_p(), ctypes.c_uint64(), ctypes.c_uint32()
    rc = L.asz_zstd_compress_mt(buf.ctypes.data_as(ctypes.c_void_p), buf.size, level, threads, ctypes.byref(out),
                                ctypes.byref(n), ctypes.byref(used))
    assert rc == 0
    try:
        return ctypes.string_at(out, n.value), used.value
    finally:
        L.asz_free(out)


def _slices(n, threads):
    # slices of whole 128 KiB blocks, none shorter than 1 MiB
    blk = 128 << 10
    sl = -(-n // threads)
    sl = max(MiB, -(-sl // blk) * blk)
    return min(threads, -(-n // sl))


def _frame_content_size(frame: bytes) -> int:
    # RFC 8878 3.1.1.1: frame header descriptor, window descriptor, FCS field
    fhd = frame[4]
    single = (fhd >> 5) & 1
    fcs_bytes = {0: 1 if single else 0, 1: 2, 2: 4, 3: 8}[fhd >> 6]
    o = 5 + (0 if single else 1) + {0: 0, 1: 1, 2: 2, 3: 4}[fhd & 3]
    v = int.from_bytes(frame[o:o + fcs_bytes], "little")
    return v + 256 if fcs_bytes == 2 else v


@pytest.mark.parametrize("size", [100, 3 * MiB + 12345, 24 * MiB + 7])
@pytest.mark.parametrize("threads", [2, 5, 16])
def test_incompressible_stream_gives_the_single_threaded_frame(size, threads):
    data = np.random.default_rng(size + threads).integers(0, 256, size, dtype=np.uint8).tobytes()
    ref = codec.zstd_compress(data)
    got, used = _mt(data, threads=threads)
    assert got == ref
    assert _frame_content_size(got) == size
    assert codec.zstd_decompress(got) == data
    assert used == _slices(size, threads)


def test_compressible_slice_falls_back_to_one_call():
    rng = np.random.default_rng(7)
    data = rng.integers(0, 256, 12 * MiB, dtype=np.uint8)
    data[6 * MiB:] = np.arange(6 * MiB, dtype=np.uint64) % 13  # a compressible second half
    data = data.tobytes()
    got, used = _mt(data, threads=4)
    assert used == 1
    assert got == codec.zstd_compress(data)
    assert codec.zstd_decompress(got) == data


def test_compressible_first_slice_is_spliced():
    # compressed blocks are allowed in slice 0 (they open the frame)
    rng = np.random.default_rng(8)
    data = rng.integers(0, 256, 8 * MiB, dtype=np.uint8)
    data[:MiB // 2] = 0
    data = data.tobytes()
    got, used = _mt(data, threads=4)
    assert used == 4
    assert len(got) < len(data)
    assert codec.zstd_decompress(got) == data


@pytest.mark.parametrize("level", [-5, 1, 3, 9])
def test_levels_decode(level):
    data = np.random.default_rng(level + 100).integers(0, 256, 5 * MiB, dtype=np.uint8).tobytes()
    for threads in (1, 3):
        got = codec.zstd_compress(data, level=level, threads=threads)
        assert codec.zstd_decompress(got) == data
