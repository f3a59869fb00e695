"""The round kernels' target -> QB slot arithmetic (asz_round_layout exposes the
same __host__ __device__ functions the kernels run) against a brute-force
replay of the reference's block_compress_impl loops (compress.tpp:51-88,
asz_param.hpp:105-129, blocks of compress.tpp:179-189).  Runs on the CPU."""
import ctypes
import itertools

import numpy as np
import pytest

from asz_amd import _lib, codec


def reference_layout(dims, plan, bpr):
    """Per round: list of (pos, slot) in block-major order, as the reference writes them."""
    D = list(dims)
    axis = [int(p[5]) for p in plan]
    order = [codec.ORDERS.index(p[7:]) for p in plan]
    ps = [1, 1, 1]
    for a in axis:
        ps[a] *= 2
    bs = [ps[a] * bpr for a in range(3)]
    bn = [(D[a] + bs[a] - 1) // bs[a] for a in range(3)]
    out = [dict() for _ in plan]
    bases = [0] * len(plan)
    for b0, b1, b2 in itertools.product(range(bn[0]), range(bn[1]), range(bn[2])):
        lo = [b0 * bs[0], b1 * bs[1], b2 * bs[2]]
        hi = [min(lo[a] + bs[a] + 1, D[a]) for a in range(3)]
        stride = list(ps)
        for r, (d, o) in enumerate(zip(axis, order)):
            first = [lo[a] + stride[a] // 2 if a == d else (0 if lo[a] == 0 else lo[a] + stride[a]) for a in range(3)]
            n = [0 if first[a] >= hi[a] else (hi[a] - first[a] + stride[a] - 1) // stride[a] for a in range(3)]
            for j0, i0 in enumerate(range(first[0], hi[0], stride[0])):
                for j1, i1 in enumerate(range(first[1], hi[1], stride[1])):
                    for j2, i2 in enumerate(range(first[2], hi[2], stride[2])):
                        if o == 0:
                            idx = (j1 * n[2] + j2) * n[0] + j0
                        elif o == 1:
                            idx = (j0 * n[2] + j2) * n[1] + j1
                        else:
                            idx = (j0 * n[1] + j1) * n[2] + j2
                        pos = (i0 * D[1] + i1) * D[2] + i2
                        assert pos not in out[r]
                        out[r][pos] = bases[r] + idx
            bases[r] += n[0] * n[1] * n[2]
            stride[d] //= 2
    return out


def kernel_layout(dims, plan, bpr, r):
    L = _lib.lib()
    L.asz_round_layout.argtypes = [ctypes.c_void_p, ctypes.POINTER(_lib.Plan), ctypes.c_uint64, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_uint64), ctypes.c_void_p, ctypes.c_void_p]
    d = np.asarray(dims, np.uint64)
    pl = codec.make_plan(plan)
    cnt = ctypes.c_uint64()
    assert L.asz_round_layout(d.ctypes.data_as(ctypes.c_void_p), ctypes.byref(pl), bpr, r, ctypes.byref(cnt),
                              None, None) == 0
    pos = np.zeros(cnt.value, np.uint64)
    slot = np.zeros(cnt.value, np.uint64)
    assert L.asz_round_layout(d.ctypes.data_as(ctypes.c_void_p), ctypes.byref(pl), bpr, r, ctypes.byref(cnt),
                              pos.ctypes.data_as(ctypes.c_void_p), slot.ctypes.data_as(ctypes.c_void_p)) == 0
    return pos, slot


CASES = [
    ((20, 23, 29), ["cubic0_012", "cubic1_021", "cubic2_120", "cubic0_120", "cubic1_012", "cubic2_021"], 3),
    ((17, 16, 33), ["cubic2_012", "cubic2_120", "cubic1_021", "cubic0_012", "cubic2_021"], 2),
    ((7, 9, 31), ["cubic0_021", "cubic2_012", "cubic2_012", "cubic1_120"], 1),
    ((33, 12, 14), ["cubic0_012", "cubic0_120", "cubic0_021", "cubic1_012", "cubic2_012"], 96),
    ((25, 26, 27), ["cubic1_120", "cubic0_021", "cubic2_012", "cubic1_012", "cubic0_012", "cubic2_120", "cubic1_021",
                    "cubic0_120", "cubic2_021"], 4),
]


@pytest.mark.parametrize("dims,plan,bpr", CASES)
def test_kernel_slots_match_reference_loops(dims, plan, bpr):
    ref = reference_layout(dims, plan, bpr)
    covered = 0
    for r in range(len(plan)):
        pos, slot = kernel_layout(dims, plan, bpr, r)
        assert len(pos) == len(ref[r])
        assert sorted(slot.tolist()) == list(range(len(pos)))  # a permutation of the round array
        got = dict(zip(pos.tolist(), slot.tolist()))
        assert got == ref[r]
        covered += len(pos)
    ps = [1, 1, 1]
    for p in plan:
        ps[int(p[5])] *= 2
    pivots = int(np.prod([(dims[a] + ps[a] - 1) // ps[a] for a in range(3)]))
    assert covered + pivots == int(np.prod(dims))
