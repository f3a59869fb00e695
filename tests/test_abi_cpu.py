"""CPU checks of the product library: it loads without a GPU, exports every
entry point include/asz_gpu.h declares, and its host-side Huffman codebook
builder reproduces the reference tree (compared with the oracle's encoder)."""
import ctypes
import os
import re

import numpy as np
import pytest

from asz_amd import _lib, codec
from oracle import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "asz_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(asz_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert sorted(_lib.EXPORTS) == names


def test_params_defaults_match_reference():
    p = codec.make_params()
    assert p.pivot_eb_ratio == np.float32(0.05) and p.eb_ratio == np.float32(0.76)
    assert p.slope_ratio == np.float32(0.54) and p.block_pivot_ratio == 4 and p.t_precision == 32
    assert p.stat_eb_ratio == np.float32(0.2) and p.h1_block_sampling_num == 256
    assert p.t_min == np.float32(1 / 16) and p.t_max == np.float32(6 / 16)


def test_geometry_matches_oracle():
    plan = ["cubic0_012", "cubic2_120", "cubic2_021", "cubic1_012", "cubic0_012"]
    ctx_free = codec.make_plan(plan)
    d = np.array([33, 40, 57], np.uint64)
    ps, bs, bn = (np.zeros(3, np.uint64) for _ in range(3))
    rc = _lib.lib().asz_geometry(d.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ctx_free), 3,
                                 ps.ctypes.data_as(ctypes.c_void_p), bs.ctypes.data_as(ctypes.c_void_p),
                                 bn.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    ops, obs, obn = api.geometry([33, 40, 57], [0, 2, 2, 1, 0], 3)
    assert list(ps) == list(ops) and list(bs) == list(obs) and list(bn) == list(obn)


def _pack(v, codes, lens, offset):
    """LSB-first packing of the codes (the reference payload layout)."""
    bits = []
    for x in v:
        s = int(x) - offset
        c, n = int(codes[s]), int(lens[s])
        bits.extend((c >> k) & 1 for k in range(n))
    out = bytearray((len(bits) + 7) // 8)
    for i, b in enumerate(bits):
        out[i >> 3] |= b << (i & 7)
    return bytes(out)


CASES = [
    ("laplace", lambda r: np.round(r.laplace(0, 3, 5000)).astype(np.int32)),
    ("wide", lambda r: np.round(r.laplace(0, 800, 3000)).astype(np.int32)),
    ("sparse_4096", lambda r: np.concatenate([np.zeros(100, np.int32), np.array([5000, -4000], np.int32)])),
    ("two", lambda r: np.array([3, 3, 3, 4], np.int32)),
    ("single", lambda r: np.full(77, -12, np.int32)),
    ("ties", lambda r: np.repeat(np.arange(-20, 20, dtype=np.int32), 5)),
    ("fib", lambda r: np.repeat(np.arange(12, dtype=np.int32), [1, 1, 2, 3, 5, 8, 13, 21, 34, 55, 89, 144])),
    ("negative_max", lambda r: -np.abs(np.round(r.laplace(0, 5, 999))).astype(np.int32) - 1),
]


@pytest.mark.parametrize("name,gen", CASES)
def test_codebook_is_reference_tree(name, gen):
    r = np.random.default_rng(3)
    v = gen(r)
    offset = int(v.min())
    hist = np.bincount(v - offset).astype(np.uint64)
    hdr, codes, lens, bits = codec.huffman_codebook(hist, offset, v.size)
    ref = api.huffman_encode_i32(v)
    if int(v.max()) == offset:
        assert hdr == ref  # single-symbol form carries the count, no payload
        return
    assert ref[:len(hdr)] == hdr
    assert ref[len(hdr):] == _pack(v, codes, lens, offset)
    assert bits == int(np.sum(hist * lens[: hist.size]))


def test_codebook_kraft_equality():
    r = np.random.default_rng(5)
    v = np.round(r.laplace(0, 20, 20000)).astype(np.int32)
    offset = int(v.min())
    hist = np.bincount(v - offset).astype(np.uint64)
    _, codes, lens, _ = codec.huffman_codebook(hist, offset, v.size)
    used = lens[hist > 0].astype(np.float64)
    assert abs(np.sum(2.0 ** -used) - 1.0) < 1e-12
    # prefix-free: no code is a prefix of another (LSB-first bit strings)
    strs = sorted(format(int(codes[i]), "064b")[::-1][: int(lens[i])] for i in np.nonzero(hist)[0])
    for a, b in zip(strs, strs[1:]):
        assert not b.startswith(a)


def test_zstd_roundtrip_host():
    payload = bytes(range(256)) * 100
    frame = codec.zstd_compress(payload)
    assert codec.zstd_decompress(frame) == payload
