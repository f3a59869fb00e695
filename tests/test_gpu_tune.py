"""The autotuner of the round kernels' tile shapes (asz_autotune) and its cache
file.  A tile shape (sweep segment length and lanes along DIM2, rows per
workgroup of the DIM2 rounds) changes only how targets map to lanes and
workgroups, so every stream must equal the default one: after autotuning,
after a save / load round trip into a fresh context, and under forced
non-default shapes for every round."""
import os
import tempfile

import numpy as np
import pytest
import torch

from asz_amd import batch, codec, fields

pytestmark = pytest.mark.gpu

PLAN = ["cubic0_012", "cubic1_012", "cubic2_012", "cubic0_021", "cubic1_120", "cubic2_012", "cubic0_012",
        "cubic1_021", "cubic2_120", "cubic0_012", "cubic1_012", "cubic2_021", "cubic0_120", "cubic1_012"]
SHAPE = (96, 112, 128)


def _compress(ctx, g, plan, params):
    w = g.clone()
    res = ctx.compress_field(w.data_ptr(), SHAPE, plan, params, tsel="gpu", copy_stream=True)
    torch.cuda.synchronize()
    return res.stream, w.cpu().numpy()


@pytest.fixture(scope="module")
def field():
    return fields.config2_field(SHAPE, seed=77, device="cuda").contiguous()


def test_autotune_keeps_the_stream(field):
    plan = codec.make_plan(PLAN)
    params = codec.make_params(rel_eb=1e-3, block_pivot_ratio=3, family="cubic")
    with codec.Context(0) as ref_ctx, codec.Context(0) as ctx, tempfile.TemporaryDirectory() as tmp:
        ref_s, ref_r = _compress(ref_ctx, field, plan, params)
        before = field.clone()
        rep = ctx.autotune(field, plan, params, reps=1)
        assert torch.equal(before, field), "the tuner must leave the field untouched"
        assert rep.rounds_tuned > 0 and rep.trials > rep.rounds_tuned
        assert rep.tuned_ms <= rep.default_ms * 1.000001
        chosen = ctx.tune_lookup(SHAPE, plan, params)
        assert chosen is not None and len(chosen) == len(PLAN)
        s, r = _compress(ctx, field, plan, params)
        assert s == ref_s and np.array_equal(r.view(np.uint32), ref_r.view(np.uint32))
        path = os.path.join(tmp, "tunes.txt")
        ctx.save_tunes(path)
        with codec.Context(0) as ctx2:
            assert ctx2.tune_lookup(SHAPE, plan, params) is None
            ctx2.load_tunes(path)
            assert ctx2.tune_lookup(SHAPE, plan, params) == chosen
            s2, r2 = _compress(ctx2, field, plan, params)
            assert s2 == ref_s and np.array_equal(r2.view(np.uint32), ref_r.view(np.uint32))


@pytest.mark.parametrize("shape_code", [0, 1, 2])
def test_forced_tile_shapes_keep_the_stream(field, shape_code):
    plan = codec.make_plan(PLAN)
    params = codec.make_params(rel_eb=1e-3, block_pivot_ratio=3, family="cubic")
    segs, xshs, rpws = (8, 12, 16), (6, 7, 8), (16, 48, 256)
    tiles = ",".join(f"{segs[(r + shape_code) % 3]}:{xshs[(r + 2 * shape_code) % 3]}:{rpws[(r + shape_code) % 3]}"
                     for r in range(len(PLAN)))
    ao = "".join(f"{p[5]}{codec.ORDERS.index(p[7:])}" for p in PLAN)
    with codec.Context(0) as ref_ctx, codec.Context(0) as ctx, tempfile.TemporaryDirectory() as tmp:
        ref_s, ref_r = _compress(ref_ctx, field, plan, params)
        path = os.path.join(tmp, "forced.txt")
        with open(path, "w") as fh:
            fh.write("asz-tune 1\n")
            fh.write(f"{SHAPE[0]} {SHAPE[1]} {SHAPE[2]} 3 {params.h1_block_sampling_num} {len(PLAN)} {ao} {tiles}\n")
        ctx.load_tunes(path)
        assert ctx.tune_lookup(SHAPE, plan, params) is not None
        s, r = _compress(ctx, field, plan, params)
        assert s == ref_s, f"forced tile shapes {tiles}: streams differ"
        assert np.array_equal(r.view(np.uint32), ref_r.view(np.uint32))


def test_tune_file_errors():
    with codec.Context(0) as ctx, tempfile.TemporaryDirectory() as tmp:
        bad = os.path.join(tmp, "bad.txt")
        for text in ("not a tune file\n", "asz-tune 1\n7 7 7 3 256 2 0212 99:6:0,0:0:0\n",
                     "asz-tune 1\n7 7 7 3 256 2 02 0:0:0\n"):
            with open(bad, "w") as fh:
                fh.write(text)
            with pytest.raises(codec.AszError):
                ctx.load_tunes(bad)


def test_batch_autotune_cache(field):
    """The batch driver tunes each (shape, plan) once and reads it back from the cache file."""
    flds = [fields.config2_field(SHAPE, seed=80 + i, device="cuda").contiguous() for i in range(3)]
    plans = [codec.make_plan(PLAN)] * 3
    with tempfile.TemporaryDirectory() as tmp:
        cache = os.path.join(tmp, "c", "tunes.txt")
        with codec.Context(0) as ctx:
            assert batch.autotune(ctx, flds, plans, cache=cache, reps=1) == 1
            assert os.path.exists(cache)
            tuned_s, _ = ctx.compress_batch([g.clone() for g in flds], plans,
                                            codec.make_params(rel_eb=1e-3, block_pivot_ratio=3, family="cubic"),
                                            tsel="grad", concurrency=2)
        with codec.Context(0) as ctx:
            assert batch.autotune(ctx, flds, plans, cache=cache, reps=1) == 0
        with codec.Context(0) as plain_ctx:
            plain_s, _ = plain_ctx.compress_batch([g.clone() for g in flds], plans,
                                                  codec.make_params(rel_eb=1e-3, block_pivot_ratio=3, family="cubic"),
                                                  tsel="grad", concurrency=2)
        assert tuned_s == plain_s
