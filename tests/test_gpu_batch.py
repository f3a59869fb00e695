"""Batched form: asz_compress_batch on fields of uneven shape (time steps
sharing one plan, mixed shapes with their own) must give, field by field,
exactly the stream the single-field path gives, and every stream decodes
within the error bound.  With one worker the five time steps go through the
multi-field kernels in chunks of three and two fields; with three workers
every chunk holds one field."""
import sys

import numpy as np
import pytest
import torch

from asz_amd import batch, codec, fields
from oracle import api

pytestmark = pytest.mark.gpu

STEP_PLAN = ["cubic0_012", "cubic1_012", "cubic2_012", "cubic2_012", "cubic1_021", "cubic0_012", "cubic2_012",
             "cubic1_012", "cubic0_120", "cubic2_012"]
MIXED = {(40, 64, 72): ["cubic2_012", "cubic1_012", "cubic0_012", "cubic2_012", "cubic1_012", "cubic2_021",
                        "cubic0_120", "cubic1_012"],
         (23, 17, 31): ["cubic0_012", "cubic2_012", "cubic1_012", "cubic2_012", "cubic0_021"]}


def _fields():
    flds = [s.contiguous() for s in fields.shifted_series((48, 56, 80), 5, seed=9, device="cuda")]
    plans = [STEP_PLAN] * len(flds)
    for k, (shp, pl) in enumerate(MIXED.items()):
        for j in range(2):
            g = fields.rescale(fields.grf(shp, -11.0 / 3.0, 70 + 2 * k + j, device="cuda"), 0.0, 1.0)
            flds.append(g.to(torch.float32).contiguous())
            plans.append(pl)
    return flds, plans


@pytest.mark.parametrize("tsel,concurrency,family", [("gpu", 3, "cubic"), ("grad", 3, "cubic"), ("gpu", 1, "cubic"),
                                                     ("grad", 1, "cubic"), ("gpu", 1, "quintic")])
def test_batch_equals_single_field_path(tsel, concurrency, family):
    flds, plans = _fields()
    originals = [f.clone() for f in flds]
    ctx = codec.Context(0)
    try:
        print("batch of", len(flds), "fields, tsel", tsel, "concurrency", concurrency, family, flush=True)
        sys.stdout.flush()
        work = [f.clone() for f in flds]
        streams, info = batch.compress_fields(ctx, work, [codec.make_plan(p) for p in plans], tsel=tsel,
                                              concurrency=concurrency, family=family)[:2]
        params = codec.make_params(rel_eb=float(np.float32(1e-3)), block_pivot_ratio=3, family=family)
        for i, (f, pl) in enumerate(zip(originals, plans)):
            w = f.clone()
            mn, mx = ctx.range(w.data_ptr(), w.numel())
            slope, *_ = ctx.default_slope(w.data_ptr(), tuple(w.shape), params, mn, mx)
            params.default_slope = slope
            assert np.float32(info[i].default_slope) == np.float32(slope)
            res = ctx.compress_device(w.data_ptr(), tuple(w.shape), codec.make_plan(pl), params, mn, mx, tsel=tsel)
            assert streams[i] == res.stream, f"field {i}: batch stream differs"
            assert torch.equal(work[i], w), f"field {i}: reconstruction differs"
            out = ctx.decompress(streams[i], tuple(w.shape), params)
            ref = f.cpu().numpy().astype(np.float64)
            eb = float(api.abs_eb(params.rel_eb, mn, mx))
            err = np.abs(out.astype(np.float64) - ref)
            assert np.all(err <= eb + np.spacing(np.abs(out)).astype(np.float64) / 2 + 1e-12 * eb)
    finally:
        ctx.close()


@pytest.mark.parametrize("bad", [2, 3])
def test_batch_error_names_the_field(bad):
    # two workers: the time steps go in chunks [0, 1], [2, 3], [4]
    flds, plans = _fields()
    flds[bad].fill_(1.0)  # a constant field: max_val == min_val
    ctx = codec.Context(0)
    try:
        with pytest.raises(codec.AszError, match=f"field {bad}: max_val==min_val"):
            batch.compress_fields(ctx, flds, [codec.make_plan(p) for p in plans], concurrency=2)
    finally:
        ctx.close()
