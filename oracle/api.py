"""ctypes front end of the oracle (test infrastructure only).

Loads _ref/liboracle.so (the CPU restatement in asz_oracle.cpp) and exposes the
reference CLI built from its own sources (_ref/s1_compress, _ref/s1_decompress).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
LIB_PATH = os.path.join(REF_DIR, "liboracle.so")
REF_COMPRESS = os.path.join(REF_DIR, "s1_compress")
REF_DECOMPRESS = os.path.join(REF_DIR, "s1_decompress")

_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built (run make -C oracle port)")
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_compress.restype = ctypes.c_int64
        L.oracle_compress.argtypes = [_P, _P, ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_float, _U64, ctypes.c_int,
                                      ctypes.c_float, _U64, ctypes.c_float, ctypes.c_float, _P, _P, _P, _P, _P]
        L.oracle_decompress.restype = ctypes.c_int
        L.oracle_decompress.argtypes = [_P, _P, _P, _P, ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_float,
                                        ctypes.c_float, ctypes.c_float, _U64, ctypes.c_int, _P]
        L.oracle_geometry.restype = ctypes.c_int
        L.oracle_geometry.argtypes = [_P, ctypes.c_int, _P, _U64, _P, _P, _P]
        L.oracle_huffman_encode_i32.restype = ctypes.c_int64
        L.oracle_huffman_encode_i32.argtypes = [_P, _U64, _P, _U64]
        L.oracle_huffman_decode_i32.restype = ctypes.c_int64
        L.oracle_huffman_decode_i32.argtypes = [_P, _U64, _P, _U64]
        L.oracle_build_payload.restype = ctypes.c_int64
        L.oracle_build_payload.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_int, _P, _P, ctypes.c_int, _P,
                                           _U64, _P, _P, _P, _U64]
        L.oracle_parse_payload.restype = ctypes.c_int64
        L.oracle_parse_payload.argtypes = [_P, _U64, _P, _P, _P, _P, _P, _P, _P, _P, _P]
        L.oracle_default_slope.restype = ctypes.c_float
        L.oracle_default_slope.argtypes = [_P, _P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                           ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                           ctypes.c_float, ctypes.c_int, _U64, _P]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _err():
    return lib().oracle_last_error().decode()


@dataclass
class Params:
    """Numeric CLI parameters (asz_param.hpp:147-183 defaults)."""
    rel_eb: float = 1e-3
    pivot_eb_ratio: float = 0.05
    eb_ratio: float = 0.76
    slope_ratio: float = 0.54
    block_pivot_ratio: int = 4
    t_precision: int = 32
    stat_eb_ratio: float = 0.2
    h1_block_sampling_num: int = 256
    default_slope: float = 1000.0


@dataclass
class Stream:
    min_val: float
    max_val: float
    axis: list
    order: list
    quintic: bool
    t: np.ndarray
    counts: np.ndarray
    qb: np.ndarray
    nbytes: int = 0
    arrays: list = field(default_factory=list)

    @property
    def L(self):
        return len(self.axis)


def f32(x):
    return float(np.float32(x))


def abs_eb(rel_eb, mn, mx):
    """abs_eb = rel_eb * (max - min) in float (s1_compress.cpp:138)."""
    return np.float32(rel_eb) * (np.float32(mx) - np.float32(mn))


def geometry(dims, axis, bpr):
    ps = np.zeros(3, np.uint64)
    bs = np.zeros(3, np.uint64)
    bn = np.zeros(3, np.uint64)
    d = np.asarray(dims, np.uint64)
    ax = np.asarray(axis, np.uint8)
    lib().oracle_geometry(_ptr(d), len(ax), _ptr(ax), bpr, _ptr(ps), _ptr(bs), _ptr(bn))
    return ps, bs, bn


def compress(data, axis, order, quintic, params: Params, t_inject=None, min_val=None, max_val=None):
    """Pivots + rounds (compress.tpp:151-231).  Returns (t_int, counts, qb, recon, fallbacks)."""
    data = np.ascontiguousarray(data, np.float32)
    dims = np.asarray(data.shape, np.uint64)
    L = len(axis)
    mn = f32(data.min()) if min_val is None else min_val
    mx = f32(data.max()) if max_val is None else max_val
    aeb = abs_eb(params.rel_eb, mn, mx)
    _, _, bn = geometry(data.shape, axis, params.block_pivot_ratio)
    nt = int(np.prod(bn)) * L
    t_out = np.zeros(nt, np.int64)
    qb = np.zeros(data.size, np.int32)
    counts = np.zeros(L + 1, np.uint64)
    recon = np.zeros_like(data)
    ax = np.asarray(axis, np.uint8)
    od = np.asarray(order, np.uint8)
    ti = None if t_inject is None else np.ascontiguousarray(t_inject, np.int64)
    fb = lib().oracle_compress(_ptr(data), _ptr(dims), L, _ptr(ax), _ptr(od), int(quintic), float(aeb),
                               params.pivot_eb_ratio, params.eb_ratio, params.slope_ratio, params.default_slope,
                               params.block_pivot_ratio, params.t_precision, params.stat_eb_ratio,
                               params.h1_block_sampling_num, mn, mx, _ptr(ti), _ptr(t_out), _ptr(qb), _ptr(counts),
                               _ptr(recon))
    if fb < 0:
        raise RuntimeError(_err())
    return t_out, counts, qb, recon, int(fb)


def default_slope(data, params: Params, eb_span=1.5, invalid_threshold=1.0, fallback=1000.0, t_min=1 / 16,
                  t_max=6 / 16):
    """get_default_slope (get_default_slope.tpp:5-83).  Returns (slope, max_invalid_ratio, fallback, reason)."""
    data = np.ascontiguousarray(data, np.float32)
    dims = np.asarray(data.shape, np.uint64)
    mn, mx = f32(data.min()), f32(data.max())
    aeb = abs_eb(params.rel_eb, mn, mx)
    info = np.zeros(3, np.float32)
    s = lib().oracle_default_slope(_ptr(data), _ptr(dims), float(aeb), params.stat_eb_ratio, eb_span,
                                   invalid_threshold, fallback, mn, mx, f32(t_min), f32(t_max), params.t_precision,
                                   params.h1_block_sampling_num, _ptr(info))
    return float(np.float32(s)), float(info[0]), bool(info[1]), ("none", "invalid_ratio", "invalid_difference")[
        int(info[2])]


def round_energy(state, axis, bpr, round_, quintic, theta):
    """FP64 residual energy per block and its gradient (oracle_round_energy)."""
    state = np.ascontiguousarray(state, np.float32)
    d = np.asarray(state.shape, np.uint64)
    ax = np.asarray(axis, np.uint8)
    th = np.ascontiguousarray(theta, np.float64)
    nb = th.size // (2 if quintic else 1)
    E = np.zeros(nb, np.float64)
    g = np.zeros(th.size, np.float64)
    L = lib()
    L.oracle_round_energy.restype = ctypes.c_int
    L.oracle_round_energy.argtypes = [_P, _P, ctypes.c_int, _P, _U64, ctypes.c_int, ctypes.c_int, _P, _P, _P]
    rc = L.oracle_round_energy(_ptr(state), _ptr(d), len(axis), _ptr(ax), bpr, round_, int(quintic), _ptr(th),
                               _ptr(E), _ptr(g))
    if rc != 0:
        raise RuntimeError(_err())
    return E, g


def decompress(stream: Stream, dims, params: Params):
    d = np.asarray(dims, np.uint64)
    recon = np.zeros(tuple(int(x) for x in dims), np.float32)
    aeb = abs_eb(params.rel_eb, stream.min_val, stream.max_val)
    ax = np.asarray(stream.axis, np.uint8)
    od = np.asarray(stream.order, np.uint8)
    rc = lib().oracle_decompress(_ptr(stream.qb), _ptr(stream.counts), _ptr(stream.t), _ptr(d), stream.L, _ptr(ax),
                                 _ptr(od), int(stream.quintic), float(aeb), params.pivot_eb_ratio, params.eb_ratio,
                                 params.block_pivot_ratio, params.t_precision, _ptr(recon))
    if rc != 0:
        raise RuntimeError(_err())
    return recon


def build_payload(min_val, max_val, axis, order, quintic, t, counts, qb):
    t = np.ascontiguousarray(t, np.int64)
    qb = np.ascontiguousarray(qb, np.int32)
    counts = np.ascontiguousarray(counts, np.uint64)
    ax = np.asarray(axis, np.uint8)
    od = np.asarray(order, np.uint8)
    cap = 64 + qb.size * 9 + t.size * 9 + 1024 * (len(axis) + 2) * 64
    out = np.zeros(cap, np.uint8)
    n = lib().oracle_build_payload(min_val, max_val, len(axis), _ptr(ax), _ptr(od), int(quintic), _ptr(t), t.size,
                                   _ptr(qb), _ptr(counts), _ptr(out), cap)
    if n < 0:
        raise RuntimeError(_err())
    return out[:n].tobytes()


def parse_payload(payload: bytes) -> Stream:
    buf = np.frombuffer(payload, np.uint8)
    mm = np.zeros(2, np.float32)
    L = ctypes.c_int()
    q = ctypes.c_int()
    nt = ctypes.c_uint64()
    ax = np.zeros(64, np.uint8)
    od = np.zeros(64, np.uint8)
    counts = np.zeros(64, np.uint64)
    n = lib().oracle_parse_payload(_ptr(buf), buf.size, _ptr(mm), ctypes.byref(L), _ptr(ax), _ptr(od),
                                   ctypes.byref(q), ctypes.byref(nt), None, _ptr(counts), None)
    if n < 0:
        raise RuntimeError(_err())
    t = np.zeros(nt.value, np.int64)
    Lv = L.value
    qb = np.zeros(int(counts[:Lv + 1].sum()), np.int32)
    n = lib().oracle_parse_payload(_ptr(buf), buf.size, _ptr(mm), ctypes.byref(L), _ptr(ax), _ptr(od),
                                   ctypes.byref(q), ctypes.byref(nt), _ptr(t), _ptr(counts), _ptr(qb))
    if n < 0:
        raise RuntimeError(_err())
    return Stream(float(mm[0]), float(mm[1]), [int(x) for x in ax[:Lv]], [int(x) for x in od[:Lv]], bool(q.value),
                  t, counts[:Lv + 1].copy(), qb, int(n))


def huffman_encode_i32(v):
    v = np.ascontiguousarray(v, np.int32)
    cap = 64 + v.size * 9 + 4096
    out = np.zeros(cap, np.uint8)
    n = lib().oracle_huffman_encode_i32(_ptr(v), v.size, _ptr(out), cap)
    if n < 0:
        raise RuntimeError(_err())
    return out[:n].tobytes()


def huffman_decode_i32(b, count):
    buf = np.frombuffer(b, np.uint8)
    out = np.zeros(count, np.int32)
    n = lib().oracle_huffman_decode_i32(_ptr(buf), buf.size, _ptr(out), count)
    if n < 0:
        raise RuntimeError(_err())
    return out, int(n)


# ---------------------------------------------------------------- reference CLI

def ref_available():
    return os.path.exists(REF_COMPRESS) and os.path.exists(REF_DECOMPRESS)


def run_ref_compress(in_path, dims, rel_eb, out_path, extra=(), timeout=600):
    cmd = [REF_COMPRESS, "-f", "-i", in_path, "-3", *map(str, dims), "-M", "REL", f"{rel_eb:.17g}", "-z", out_path,
           *map(str, extra)]
    return subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, check=True).stdout


def run_ref_decompress(asz_path, dims, rel_eb, out_path, extra=(), timeout=600):
    cmd = [REF_DECOMPRESS, "-f", "-z", asz_path, "-3", *map(str, dims), "-M", "REL", f"{rel_eb:.17g}", "-o",
           out_path, *map(str, extra)]
    return subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, check=True).stdout
