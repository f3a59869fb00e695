"""Block rounds against the separate t-selection and round launches.

The leading rounds with at most 16384 targets per block run as block rounds:
one launch per round in which a workgroup selects its block's t and then
predicts the block's targets (coarse.hip).  Whatever the number of rounds
that run this way (ASZ_OPT_COARSE_ROUNDS: every qualifying round, a few, or
none), every call must give the same t array, QB arrays, reconstruction and
stream; both follow the block-major loop of compress.tpp:179-208.  Covered:
the injected t (the reference's t), the five-candidate scan, the
gradient-tuned form, the exact host scoring, the quintic predictor, many
small blocks (block pivot ratio 1), a single block, and the multi-field
batch path."""
import numpy as np
import pytest
import torch

from asz_amd import codec, fields

pytestmark = pytest.mark.gpu

PLAN = ["cubic0_012", "cubic1_012", "cubic2_012", "cubic0_021", "cubic1_120", "cubic2_012", "cubic0_012",
        "cubic1_021", "cubic2_120", "cubic0_012", "cubic1_012", "cubic2_021", "cubic0_120", "cubic1_012"]


@pytest.fixture(scope="module")
def ctx():
    c = codec.Context(0)
    yield c
    c.set_coarse_rounds(-1)
    c.close()


def _field(shape, seed):
    return fields.to_numpy(fields.config2_field(shape, seed=seed)).astype(np.float32)


def _run(ctx, f, plan, params, tsel, coarse, t_inject=None):
    ctx.set_coarse_rounds(coarse)
    res = ctx.compress(f, codec.make_plan(plan), params, tsel=tsel, t_inject=t_inject, want_qb=True, want_t=True,
                       want_recon=True)
    ctx.set_coarse_rounds(-1)
    return res


@pytest.mark.parametrize("shape,plan_len,bpr,fam,tsel,rel", [
    ((48, 56, 80), 14, 3, "cubic", "gpu", 1e-3),
    ((48, 56, 80), 14, 3, "quintic", "gpu", 1e-4),
    ((48, 56, 80), 14, 3, "cubic", "grad", 1e-3),
    ((40, 44, 52), 14, 1, "cubic", "gpu", 1e-3),     # many small blocks
    ((33, 29, 31), 10, 8, "cubic", "gpu", 1e-2),     # one block
    ((64, 96, 72), 14, 2, "quintic", "grad", 1e-3),
    ((40, 48, 56), 12, 3, "cubic", "exact", 1e-3),
])
def test_coarse_equals_per_round(ctx, shape, plan_len, bpr, fam, tsel, rel):
    print(f"coarse vs per-round: shape {shape} bpr {bpr} {fam} tsel {tsel} rel {rel}", flush=True)
    f = _field(shape, seed=sum(shape))
    plan = PLAN[:plan_len]
    params = codec.make_params(rel_eb=rel, block_pivot_ratio=bpr, family=fam)
    ref = _run(ctx, f, plan, params, tsel, 0)
    assert int(ref.stats.coarse_rounds) == 0
    for coarse in (-1, 3):
        got = _run(ctx, f, plan, params, tsel, coarse)
        assert int(got.stats.coarse_rounds) > 0, "no round ran as a block round"
        if coarse == 3:
            assert int(got.stats.coarse_rounds) <= 3
        assert np.array_equal(got.t, ref.t), f"t arrays differ (coarse {coarse})"
        assert np.array_equal(got.qb, ref.qb), f"QB arrays differ (coarse {coarse})"
        assert np.array_equal(got.recon.view(np.uint32), ref.recon.view(np.uint32)), "reconstructions differ"
        assert got.stream == ref.stream, f"streams differ (coarse {coarse})"
        # the stream decodes to the same reconstruction
        out = ctx.decompress(got.stream, f.shape, params)
        assert np.array_equal(out.view(np.uint32), ref.recon.view(np.uint32))


def test_coarse_with_injected_t(ctx):
    """The reference path of the parity tests: t given, every round's QBs equal."""
    shape = (48, 56, 80)
    f = _field(shape, seed=5)
    params = codec.make_params(rel_eb=1e-3, block_pivot_ratio=3, family="cubic")
    base = _run(ctx, f, PLAN, params, "gpu", 0)
    rng = np.random.default_rng(3)
    t_inj = rng.choice(np.array([2, 8, 9, 10, 11], np.int64), size=base.t.shape)  # candidate t x 32
    a = _run(ctx, f, PLAN, params, "inject", 0, t_inject=t_inj)
    b = _run(ctx, f, PLAN, params, "inject", -1, t_inject=t_inj)
    assert int(b.stats.coarse_rounds) > 0
    assert np.array_equal(a.qb, b.qb)
    assert np.array_equal(a.recon.view(np.uint32), b.recon.view(np.uint32))
    assert a.stream == b.stream


def test_coarse_in_batch(ctx):
    """Multi-field chunks: items (round, field, block) of every field in one launch."""
    shapes = [(48, 56, 80)] * 5 + [(40, 44, 52)] * 3
    flds = [fields.config2_field(s, seed=40 + i, device="cuda").contiguous() for i, s in enumerate(shapes)]
    plans = [codec.make_plan(PLAN) for _ in shapes]
    params = codec.make_params(rel_eb=1e-3, block_pivot_ratio=3, family="cubic")
    outs = {}
    for coarse in (0, -1):
        ctx.set_coarse_rounds(coarse)
        work = [g.clone() for g in flds]
        streams, _ = ctx.compress_batch(work, plans, params, tsel="grad", concurrency=3)
        torch.cuda.synchronize()
        outs[coarse] = (streams, [w.cpu().numpy() for w in work])
    ctx.set_coarse_rounds(-1)
    for i in range(len(shapes)):
        assert outs[0][0][i] == outs[-1][0][i], f"field {i}: streams differ"
        assert np.array_equal(outs[0][1][i].view(np.uint32), outs[-1][1][i].view(np.uint32))


@pytest.mark.parametrize("tsel", ["gpu", "grad"])
def test_all_block_rounds_match_oracle(ctx, tsel):
    """Every round a block round (one block, the last round included: its t
    values come from its own launch, before the t array is histogrammed): the
    stream is the oracle's for the t the GPU chose."""
    from oracle import api
    shape = (33, 29, 31)
    f = _field(shape, seed=11)
    plan = PLAN[:10]
    params = codec.make_params(rel_eb=1e-2, block_pivot_ratio=8, family="cubic")
    res = ctx.compress(f, codec.make_plan(plan), params, tsel=tsel, want_t=True)
    assert int(res.stats.coarse_rounds) == len(plan)
    axis = [int(s[5]) for s in plan]
    order = [codec.ORDERS.index(s[7:]) for s in plan]
    op = api.Params(rel_eb=float(np.float32(1e-2)), block_pivot_ratio=8)
    t, counts, qb, recon, _ = api.compress(f, axis, order, False, op, t_inject=res.t, min_val=res.min_val,
                                           max_val=res.max_val)
    assert res.stream == api.build_payload(res.min_val, res.max_val, axis, order, False, t, counts, qb)
