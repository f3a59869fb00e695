"""The allocation-free range scan (asz_range_f32_noalloc) against asz_range_f32,
through the C ABI: the same min / max bits and the same errors, call after
call on one GPU (the scratch array keeps the last call's words, so every call
must start from its own preset), the second field smaller than the first."""
import ctypes
import sys

import numpy as np
import pytest
import torch

from asz_amd import _lib, codec, fields

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _scan(ctx, fn, t):
    mn, mx = ctypes.c_float(), ctypes.c_float()
    ctx.after_torch()
    rc = fn(ctx._c, ctypes.c_void_p(t.data_ptr()), t.numel(), ctypes.byref(mn), ctypes.byref(mx))
    msg = _lib.lib().asz_last_error(ctx._c).decode() if rc else ""
    return rc, np.float32(mn.value).view(np.uint32), np.float32(mx.value).view(np.uint32), msg


def _field(shape, seed, kind):
    f = fields.to_numpy(fields.rescale(fields.grf(shape, -11.0 / 3.0, seed), -3.0, 5.0)).astype(np.float32)
    flat = f.reshape(-1)
    if kind == "zero_ties":
        # the smallest values are DAZ zeros (denormals and signed zeros): the
        # first of them is the minimum, with its own bits
        f = np.abs(f)
        flat = f.reshape(-1)
        flat[[5, 77 % flat.size, flat.size - 3]] = np.array([1e-40, -0.0, 0.0], np.float32)
        flat[flat < 1e-3] = np.float32(1e-39)
    elif kind == "nan":
        flat[len(flat) // 3] = np.float32("nan")
    elif kind == "constant":
        f[...] = np.float32(2.5)
    return torch.from_numpy(f).cuda() if torch.cuda.is_available() else f


CASES = [((64, 80, 96), 1, "grf"), ((7, 9, 11), 2, "grf"), ((40, 50, 60), 3, "zero_ties"), ((9, 8, 7), 4, "zero_ties"),
         ((20, 30, 40), 5, "nan"), ((10, 10, 10), 6, "constant"), ((30, 31, 33), 7, "grf")]


def test_noalloc_range_equals_range():
    L = _lib.lib()
    with codec.Context(0) as ctx:
        for shape, seed, kind in CASES:
            _say("no-alloc range case", shape, seed, kind)
            t = _field(shape, seed, kind)
            want = _scan(ctx, L.asz_range_f32, t)
            got = _scan(ctx, L.asz_range_f32_noalloc, t)
            assert got == want, (shape, kind, got, want)
            assert got[0] == (-2 if kind in ("nan", "constant") else 0)  # ASZ_ERR_DATA / ASZ_OK


def test_noalloc_range_on_two_contexts_in_turn():
    # both contexts of one GPU share the array; calls in turn (not overlapping) agree
    L = _lib.lib()
    a = _field((48, 40, 36), 11, "grf")
    b = _field((12, 14, 16), 12, "zero_ties")
    with codec.Context(0) as c1, codec.Context(0) as c2:
        for ctx, t in ((c1, a), (c2, b), (c1, b), (c2, a)):
            assert _scan(ctx, L.asz_range_f32_noalloc, t) == _scan(ctx, L.asz_range_f32, t)
