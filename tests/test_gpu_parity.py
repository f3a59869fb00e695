"""GPU parity: the HIP path (through the C ABI) against the reference's own
streams (golden fixtures) and against the oracle on generated fields.

Bar: bit-exact QB arrays, t arrays, pre-Zstd streams and reconstructions.
The GPU t scorer (ASZ_TSEL_GPU) sums in a different order than the
reference's unordered_map, so its t agreement is reported and bounded, and
every stream it writes is checked bit-exactly against the oracle given the
same t.
"""
import hashlib
import sys

import numpy as np
import pytest

from asz_amd import codec, fields
from oracle import api
from tests import golden_util as gu

pytestmark = pytest.mark.gpu

NAMES = gu.names()


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available()
    c = codec.Context(0)
    yield c
    c.close()


def _params(info, slope=1000.0):
    s = gu.settings(info)
    fam = info["family"]
    return codec.make_params(rel_eb=s["rel_eb"], pivot_eb_ratio=s["pivot_eb_ratio"], eb_ratio=s["eb_ratio"],
                             slope_ratio=s["slope_ratio"], block_pivot_ratio=s["block_pivot_ratio"],
                             t_precision=s["t_precision"], stat_eb_ratio=s["stat_eb_ratio"],
                             h1_block_sampling_num=s["h1_block_sampling_num"], t_min=s["t_min"], t_max=s["t_max"],
                             default_slope=slope, family=fam)


def _oparams(p):
    return api.Params(rel_eb=p.rel_eb, pivot_eb_ratio=p.pivot_eb_ratio, eb_ratio=p.eb_ratio,
                      slope_ratio=p.slope_ratio, block_pivot_ratio=p.block_pivot_ratio, t_precision=p.t_precision,
                      stat_eb_ratio=p.stat_eb_ratio, h1_block_sampling_num=p.h1_block_sampling_num,
                      default_slope=p.default_slope)


def _slope_on_device(ctx, data, params):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(data)).cuda()
    torch.cuda.synchronize()
    mn, mx = ctx.range(t.data_ptr(), t.numel())
    return ctx.default_slope(t.data_ptr(), data.shape, params, mn, mx), (mn, mx)


@pytest.mark.parametrize("name", NAMES)
def test_inject_t_stream_is_byte_identical(ctx, name):
    info, data, asz = gu.load(name)
    _say("case", name, data.shape)
    payload = codec.zstd_decompress(asz)
    st = api.parse_payload(payload)
    p = _params(info)
    res = ctx.compress(data, codec.make_plan(codec.plan_strings(st.axis, st.order)), p, tsel="inject",
                       t_inject=st.t, want_qb=True, want_t=True, want_recon=True)
    assert np.array_equal(res.t, st.t)
    assert np.array_equal(res.qb, st.qb)
    assert hashlib.sha256(res.recon.tobytes()).hexdigest() == info["recon_sha256"]
    assert res.stream == payload


@pytest.mark.parametrize("name", NAMES)
def test_wide_code_encoder_path_writes_the_reference_stream(ctx, name):
    """The encoder's path for codes longer than 32 bits (taken here whatever
    the longest code is) writes the same bytes as its default path."""
    info, data, asz = gu.load(name)
    _say("case", name, data.shape, "wide codes")
    payload = codec.zstd_decompress(asz)
    st = api.parse_payload(payload)
    p = _params(info)
    ctx.set_wide_codes(True)
    try:
        res = ctx.compress(data, codec.make_plan(codec.plan_strings(st.axis, st.order)), p, tsel="inject",
                           t_inject=st.t)
    finally:
        ctx.set_wide_codes(False)
    assert res.stream == payload


def test_wide_code_encoder_path_on_a_generated_field(ctx):
    """Both encoder paths on a field of many chunks (192 x 160 x 224, every
    array cut into several 4096-symbol chunks): the same stream."""
    shape = (192, 160, 224)
    data = fields.to_numpy(fields.config2_field(shape, seed=5)).astype(np.float32)
    plan = codec.make_plan(["cubic0_012", "cubic1_120", "cubic2_021", "cubic0_012", "cubic1_012", "cubic2_012",
                            "cubic0_120", "cubic1_021", "cubic2_012", "cubic0_012", "cubic1_012", "cubic2_012"])
    p = codec.make_params(rel_eb=1e-4, block_pivot_ratio=3, family="cubic", default_slope=1000.0)
    _say("case generated", shape, "default encoder path")
    base = ctx.compress(data, plan, p, tsel="gpu")
    ctx.set_wide_codes(True)
    try:
        _say("case generated", shape, "wide-code encoder path")
        wide = ctx.compress(data, plan, p, tsel="gpu")
    finally:
        ctx.set_wide_codes(False)
    assert len(base.stream) > 64 * 4096
    assert wide.stream == base.stream


@pytest.mark.parametrize("name", NAMES)
def test_gpu_decodes_reference_stream(ctx, name):
    info, data, asz = gu.load(name)
    _say("case", name, data.shape)
    payload = codec.zstd_decompress(asz)
    st = api.parse_payload(payload)
    out, qb, t = ctx.decompress(payload, data.shape, _params(info), want_qb=True, want_t=True, nt=st.t.size)
    assert np.array_equal(t, st.t)
    assert np.array_equal(qb, st.qb)
    assert hashlib.sha256(out.tobytes()).hexdigest() == info["recon_sha256"]


@pytest.mark.parametrize("name", NAMES)
def test_default_slope_and_exact_t_selection(ctx, name):
    info, data, asz = gu.load(name)
    _say("case", name, data.shape)
    payload = codec.zstd_decompress(asz)
    st = api.parse_payload(payload)
    p = _params(info)
    (slope, inv, fb, reason), _ = _slope_on_device(ctx, data, p)
    line = [l for l in info["compress_stdout"] if l.startswith("Default slope statistics")][0]
    assert line == (f"Default slope statistics: max_invalid_ratio={inv:.6f}, fallback={'true' if fb else 'false'}, "
                    f"reason={reason}, slope={slope:.6f}")
    p.default_slope = slope
    res = ctx.compress(data, codec.make_plan(codec.plan_strings(st.axis, st.order)), p, tsel="exact", want_t=True)
    assert np.array_equal(res.t, st.t)
    assert res.stream == payload


def test_gpu_t_selection_agreement(ctx):
    agree = total = 0
    for name in NAMES:
        info, data, asz = gu.load(name)
        st = api.parse_payload(codec.zstd_decompress(asz))
        p = _params(info)
        (slope, *_), _ = _slope_on_device(ctx, data, p)
        p.default_slope = slope
        res = ctx.compress(data, codec.make_plan(codec.plan_strings(st.axis, st.order)), p, tsel="gpu", want_t=True)
        agree += int(np.sum(res.t == st.t))
        total += st.t.size
        # whatever t the GPU chose, the stream is the oracle's for that t
        t, counts, qb, recon, _ = api.compress(data, st.axis, st.order, st.quintic, _oparams(p), t_inject=res.t,
                                               min_val=st.min_val, max_val=st.max_val)
        assert res.stream == api.build_payload(st.min_val, st.max_val, st.axis, st.order, st.quintic, t, counts, qb)
    _say("GPU t agreement with the reference:", agree, "/", total)
    assert agree >= 0.97 * total


PLANS = {
    "mixed14": ["cubic0_012", "cubic1_012", "cubic2_012", "cubic0_021", "cubic1_120", "cubic2_012", "cubic0_012",
                "cubic1_021", "cubic2_120", "cubic0_012", "cubic1_012", "cubic2_021", "cubic0_120", "cubic1_012"],
    "z_last": ["cubic2_012", "cubic1_012", "cubic0_012", "cubic0_120", "cubic1_021", "cubic2_012", "cubic1_120",
               "cubic2_021", "cubic0_012", "cubic1_012", "cubic2_012"],
    # the largest rounds in orders whose QBs are staged in "012" order and
    # moved into place: axis 2 / "120" (row kernel) and axis 1 / "120"
    # (sweep), both above the staging threshold at 96x128x160; axis 0 /
    # "021" below it
    "staged14": ["cubic0_012", "cubic1_012", "cubic2_012", "cubic0_012", "cubic1_012", "cubic2_012", "cubic0_012",
                 "cubic1_012", "cubic2_012", "cubic0_021", "cubic1_120", "cubic0_021", "cubic2_120", "cubic1_120"],
}


@pytest.mark.parametrize("shape,rel,plan,fam,bpr", [
    ((96, 128, 160), 1e-3, "mixed14", "cubic", 3),
    ((96, 128, 160), 1e-2, "mixed14", "quintic", 3),
    ((128, 96, 100), 1e-4, "z_last", "quintic", 96),
    ((70, 71, 90), 1e-5, "z_last", "cubic", 96),
    ((96, 128, 160), 1e-3, "staged14", "cubic", 3),
    ((96, 128, 160), 1e-4, "staged14", "quintic", 3),
])
def test_generated_fields_against_oracle(ctx, shape, rel, plan, fam, bpr):
    _say("case", shape, rel, plan, fam)
    f = fields.to_numpy(fields.config2_field(shape, seed=11)).astype(np.float32)
    p = codec.make_params(rel_eb=float(np.float32(rel)), block_pivot_ratio=bpr, family=fam)
    (slope, *_), _ = _slope_on_device(ctx, f, p)
    p.default_slope = slope
    pl = PLANS[plan]
    res = ctx.compress(f, codec.make_plan(pl), p, tsel="gpu", want_qb=True, want_t=True, want_recon=True)
    axis = [int(x[5]) for x in pl]
    order = [codec.ORDERS.index(x[7:]) for x in pl]
    t, counts, qb, recon, _ = api.compress(f, axis, order, fam == "quintic", _oparams(p), t_inject=res.t)
    assert np.array_equal(res.qb, qb)
    assert np.array_equal(res.recon.view(np.uint32), recon.view(np.uint32))
    assert res.stream == api.build_payload(res.min_val, res.max_val, axis, order, fam == "quintic", t, counts, qb)
    out = ctx.decompress(res.stream, f.shape, p)
    assert np.array_equal(out.view(np.uint32), recon.view(np.uint32))
    eb = float(api.abs_eb(p.rel_eb, res.min_val, res.max_val))
    err = np.abs(f.astype(np.float64) - out.astype(np.float64))
    assert np.all(err <= eb + np.spacing(np.abs(out)).astype(np.float64) / 2 + 1e-12 * eb)


def test_errors_are_reported(ctx):
    f = np.ones((8, 8, 8), np.float32)
    p = codec.make_params(rel_eb=1e-3, block_pivot_ratio=3)
    with pytest.raises(codec.AszError, match="max_val==min_val"):
        ctx.compress(f, codec.make_plan(["cubic0_012"]), p)
    f[1, 2, 3] = np.inf
    with pytest.raises(codec.AszError, match="nonfinite"):
        ctx.compress(f, codec.make_plan(["cubic0_012"]), p)
    info, data, asz = gu.load("small_9x13x17")
    payload = codec.zstd_decompress(asz)
    with pytest.raises(codec.AszError):
        ctx.decompress(payload[: len(payload) // 2], data.shape, _params(info))
