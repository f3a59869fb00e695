"""Parity at the full sizes of the large GPU workloads.

* 1024^3 exp(|k|^-3 GRF) density-like field in [0, 1e3] at REL 1e-4, block
  pivot ratio 96 (about 2^16 distinct QB symbols: the reference's sparse
  Huffman path, huffman.hpp:278).  The reference s1_compress (built from its
  own sources into oracle/_ref) compresses the field with the plan it picked
  for it; the GPU compressor, given that plan and the reference's per-block t,
  must write the reference's pre-Zstd stream byte for byte
  (compress.tpp:146-321), and the GPU decoder must turn the reference stream
  into s1_decompress's output bit for bit (decompress.tpp:78-208).
* The 60-field batch: 48 time steps of 100x500x500 (a |k|^-11/3 field shifted
  0.5 voxel per step along axis 2 with 5 % fresh noise) sharing step 0's saved
  plan, and 12 mixed-shape fields with their own plans, at REL 1e-3.  With the
  gradient-tuned t every field's batched stream and reconstruction equal the
  single-field call's, every stream decodes within the error bound, and the
  total compressed size stays within 1 % of the plain per-field form's
  (discrete t scan).
* The 512^3 half-smooth / half-rough field at REL 1e-3 (block pivot ratio 3)
  and 1e-4 (96), with the plans the reference picked for it: the FP32 adjoint
  of the finest round's residual energy against the FP64 CPU gradient and
  against central differences (h = 1e-3) of the FP64 energy, 1e-4 relative
  (against the Cauchy-Schwarz scale of the summands where a block's gradient
  is nearly zero).
"""
import os
import sys
import tempfile

import numpy as np
import pytest
import torch

from asz_amd import batch, codec, fields, tune
from oracle import api

pytestmark = pytest.mark.gpu

# The plan the reference s1_compress picked for the 1024^3 field (auto plan
# generation, cubic family, block pivot ratio 96, REL 1e-4).
PLAN_C3 = ["cubic1_012", "cubic0_012", "cubic0_012", "cubic0_012", "cubic1_012", "cubic2_012", "cubic2_012",
           "cubic1_012", "cubic1_012", "cubic2_012", "cubic0_120", "cubic0_120", "cubic1_120", "cubic2_120"]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _within_bound(orig: torch.Tensor, out: torch.Tensor, eb: float):
    """|x - x'| <= eb + ulp(x')/2 (the reference rounds its double
    reconstruction to float), checked on the GPU slab by slab."""
    worst = 0.0
    for s in range(0, orig.shape[0], 64):
        o = orig[s:s + 64].double()
        r = out[s:s + 64]
        spacing = (torch.nextafter(r.abs(), torch.full_like(r, float("inf"))) - r.abs()).double()
        err = (o - r.double()).abs()
        assert bool((err <= eb + spacing / 2 + 1e-12 * eb).all()), f"error bound broken in slab {s}"
        worst = max(worst, float(err.max()) / eb)
    return worst


@pytest.mark.timeout(900)
def test_1024cube_reference_stream_and_decode():
    if not api.ref_available():
        pytest.skip("reference CLI not built (oracle/_ref)")
    shape, rel, bpr = (1024, 1024, 1024), 1e-4, 96
    _say("case 1024^3 exp-GRF REL 1e-4: field")
    g = fields.config3_field(shape, seed=3, device="cuda").contiguous()
    f = g.cpu().numpy()
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        path, asz, rec = (os.path.join(tmp, n) for n in ("c3.f32", "c3.asz", "c3.rec"))
        f.tofile(path)
        common = ["--block-pivot-ratio", bpr, "--predictor-family", "cubic"]
        _say("  reference s1_compress")
        api.run_ref_compress(path, shape, rel, asz, common + ["--predictor-plan", ",".join(PLAN_C3)])
        _say("  reference s1_decompress")
        api.run_ref_decompress(asz, shape, rel, rec, common)
        os.remove(path)
        with open(asz, "rb") as fh:
            payload = codec.zstd_decompress(fh.read())
        st = api.parse_payload(payload)
        assert codec.plan_strings(st.axis, st.order) == PLAN_C3
        # the sparse-map Huffman path of the reference: a wide last-round alphabet
        p = codec.make_params(rel_eb=float(np.float32(rel)), block_pivot_ratio=bpr, family="cubic")
        _say("  GPU compress, reference t injected")
        res = ctx_compress_inject(f, PLAN_C3, p, st.t)
        assert len(res.stream) == len(payload)
        assert res.stream == payload, "GPU stream differs from the reference stream"
        del f
        _say("  GPU decode of the reference stream")
        ref_out = np.fromfile(rec, np.float32).reshape(shape)
    with codec.Context(0) as ctx:
        out = ctx.decompress(payload, shape, p)
    assert np.array_equal(out.view(np.uint32), ref_out.view(np.uint32)), "GPU decode differs from s1_decompress"
    del ref_out
    eb = float(api.abs_eb(p.rel_eb, res.min_val, res.max_val))
    worst = _within_bound(g, torch.from_numpy(out).cuda(), eb)
    _say("  stream bytes", len(payload), "max|err|/eb", worst)


def ctx_compress_inject(f, plan, p, t):
    with codec.Context(0) as ctx:
        return ctx.compress(f, codec.make_plan(plan), p, tsel="inject", t_inject=t)


@pytest.mark.timeout(900)
def test_60_field_batch_equals_single_field_calls():
    flds = batch.c4_fields()
    plans = batch.c4_plans(flds)
    assert len(flds) == 60
    assert sum(f.numel() for f in flds) == 1_463_619_660
    params = codec.make_params(rel_eb=float(np.float32(1e-3)), block_pivot_ratio=3, family="cubic")
    with codec.Context(0) as ctx:
        _say("case 60-field batch, gradient-tuned t, 6 workers")
        work = [f.clone() for f in flds]
        streams, info = batch.compress_fields(ctx, work, plans, tsel="grad", concurrency=6)[:2]
        tuned_bytes = plain_bytes = 0
        worst = 0.0
        for i, (f, pl) in enumerate(zip(flds, plans)):
            if i % 12 == 0:
                _say("  field", i, tuple(f.shape))
            shape = tuple(f.shape)
            w = f.clone()
            mn, mx = ctx.range(w.data_ptr(), w.numel())
            slope, *_ = ctx.default_slope(w.data_ptr(), shape, params, mn, mx)
            params.default_slope = slope
            res = ctx.compress_device(w.data_ptr(), shape, pl, params, mn, mx, tsel="grad")
            assert streams[i] == res.stream, f"field {i}: batched stream differs from the single-field call"
            assert torch.equal(work[i].view(torch.int32), w.view(torch.int32)), f"field {i}: reconstruction differs"
            out = torch.from_numpy(ctx.decompress(streams[i], shape, params)).cuda()
            assert torch.equal(out.view(torch.int32), w.view(torch.int32)), f"field {i}: decode differs"
            eb = float(api.abs_eb(params.rel_eb, mn, mx))
            worst = max(worst, _within_bound(f, out, eb))
            w2 = f.clone()
            plain = ctx.compress_device(w2.data_ptr(), shape, pl, params, mn, mx, tsel="gpu")
            tuned_bytes += len(codec.zstd_compress(streams[i]))
            plain_bytes += len(codec.zstd_compress(plain.stream))
        _say("  compressed bytes: tuned", tuned_bytes, "plain", plain_bytes, "max|err|/eb", worst)
        assert tuned_bytes <= 1.01 * plain_bytes


@pytest.mark.timeout(900)
@pytest.mark.parametrize("rel", [1e-3, 1e-4])
def test_512cube_half_rough_adjoint(rel):
    shape = (512, 512, 512)
    bpr, plan = tune.C5_PLANS[rel]
    axis = [int(s[5]) for s in plan]
    _say("case 512^3 half-smooth / half-rough REL", rel)
    f = fields.to_numpy(fields.config5_field(shape, seed=5, device="cuda")).astype(np.float32)
    p = codec.make_params(rel_eb=float(np.float32(rel)), block_pivot_ratio=bpr, family="cubic")
    with codec.Context(0) as ctx:
        res = ctx.compress(f, codec.make_plan(plan), p, tsel="gpu", want_t=True)
        r = len(plan) - 1  # the finest round
        st = tune.RoundState(ctx, f, codec.make_plan(plan), p, r, res.t)
        state = st.work.cpu().numpy()
        _say("  blocks", st.nblocks, "round", r)
        t = np.full(st.nblocks, 0.25)
        E64, g64 = api.round_energy(state, axis, bpr, r, False, t)
        _, g64b = api.round_energy(state, axis, bpr, r, False, t + 0.05)
        H64 = (g64b - g64) / 0.05
        theta = torch.full((st.nblocks,), 0.25, device="cuda")
        E = st.energy(theta).cpu().numpy()
        assert np.allclose(E, E64, rtol=1e-5, atol=1e-12 * E64.max())
        g32 = st.adjoint(theta, torch.ones(st.nblocks, device="cuda")).cpu().numpy()
        scale = 2.0 * np.sqrt(np.maximum(E64, 0) * np.maximum(H64, 0) / 2.0)
        ref = np.maximum(np.abs(g64), scale)
        assert np.all(np.abs(g32 - g64) <= 1e-4 * ref + 1e-30), "adjoint differs from the FP64 gradient"
        h = 1e-3
        Ep, _ = api.round_energy(state, axis, bpr, r, False, t + h)
        Em, _ = api.round_energy(state, axis, bpr, r, False, t - h)
        fd = (Ep - Em) / (2 * h)
        assert np.all(np.abs(g32 - fd) <= 1e-4 * ref + 1e-30), "adjoint differs from central differences"
        _say("  max rel. error vs FP64", float(np.max(np.abs(g32 - g64) / (ref + 1e-300))))
