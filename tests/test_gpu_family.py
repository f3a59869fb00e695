"""Predictor-family selection on the GPU (asz_select_family) against the
reference's select_predictor_family (predictor_family.tpp:217-251).

Bar: the reference's whole "Predictor family selection: ..." line, character
for character: the sampled blocks and budgets, the point count, and the
long double MSEs, PSNRs, H1s and score gain printed with %.12Le (so the
GPU probe compressions, the host t choices and the x87 sums and logarithms
agree with the reference's to the last printed digit), and the family.
* golden fixtures compressed without --predictor-family (the line is in the
  fixture), with the plan the reference used;
* generated fields with eight sample blocks, smooth and rough, fine and
  coarse error bounds (so both probe block pivot ratios and both families
  occur), against the reference CLI built from its sources (oracle/_ref).
"""
import os
import re
import sys
import tempfile

import numpy as np
import pytest

from asz_amd import codec, fields
from oracle import api
from tests import golden_util as gu

pytestmark = pytest.mark.gpu

NAMES = [n for n in gu.names()
         if any(l.startswith("Predictor family selection: ") for l in gu.load(n)[0]["compress_stdout"])]


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


def _select(ctx, data, settings, plan):
    import torch
    p = codec.make_params(rel_eb=settings["rel_eb"], pivot_eb_ratio=settings["pivot_eb_ratio"],
                          eb_ratio=settings["eb_ratio"], slope_ratio=settings["slope_ratio"],
                          block_pivot_ratio=settings["block_pivot_ratio"], t_precision=settings["t_precision"],
                          stat_eb_ratio=settings["stat_eb_ratio"],
                          h1_block_sampling_num=settings["h1_block_sampling_num"], t_min=settings["t_min"],
                          t_max=settings["t_max"], family="cubic")
    t = torch.from_numpy(np.ascontiguousarray(data)).cuda()
    torch.cuda.synchronize()
    mn, mx = ctx.range(t.data_ptr(), t.numel())
    p.default_slope = ctx.default_slope(t.data_ptr(), data.shape, p, mn, mx)[0]
    rep = ctx.select_family(t.data_ptr(), data.shape, codec.make_plan(plan), p, mn, mx)
    return codec.family_line(rep), rep


@pytest.mark.parametrize("name", NAMES)
def test_family_line_matches_fixture(ctx, name):
    info, data, _ = gu.load(name)
    _say("case", name, data.shape)
    line, rep = _select(ctx, data, gu.settings(info), info["plan"])
    want = [l for l in info["compress_stdout"] if l.startswith("Predictor family selection: ")][0]
    assert line == want
    assert ("quintic" if rep.quintic else "cubic") == info["family"]


def _grf(shape, exponent, seed, stretch=(1, 1, 1)):
    f = fields.rescale(fields.grf(shape, exponent, seed, stretch=stretch), -80.0, 40.0)
    return fields.to_numpy(f).astype(np.float32)


PLAN_A = ["cubic0_012", "cubic1_012", "cubic2_012", "cubic0_012", "cubic1_120", "cubic2_021", "cubic0_012",
          "cubic1_012", "cubic2_012", "cubic0_120", "cubic1_021", "cubic2_012"]
PLAN_B = ["cubic2_012", "cubic1_012", "cubic0_012", "cubic2_012", "cubic1_012", "cubic0_012", "cubic2_120",
          "cubic1_021", "cubic0_012", "cubic2_012"]

GENERATED = [
    ("smooth_96x112x128_rel1e-3", (96, 112, 128), 1e-3, PLAN_A, lambda s: _grf(s, -5.0, 31)),
    ("smooth_96x112x128_rel1e-5", (96, 112, 128), 1e-5, PLAN_A, lambda s: _grf(s, -5.0, 31)),
    ("turb_80x120x100_rel1e-3", (80, 120, 100), 1e-3, PLAN_B, lambda s: _grf(s, -11 / 3, 32, stretch=(4, 1, 2))),
    ("rough_64x100x90_rel1e-4", (64, 100, 90), 1e-4, PLAN_B, lambda s: _grf(s, -5 / 3, 33)),
    ("config1_100x500x500", (100, 500, 500), 1e-3, PLAN_A,
     lambda s: fields.to_numpy(fields.config1_field(s)).astype(np.float32)),
]


@pytest.mark.parametrize("name,shape,rel,plan,gen", GENERATED, ids=[g[0] for g in GENERATED])
def test_family_line_matches_reference_binary(ctx, name, shape, rel, plan, gen):
    if not os.path.exists(api.REF_COMPRESS):
        pytest.skip("reference CLI not built (oracle/Makefile ref)")
    _say("case", name, shape, rel)
    data = gen(shape)
    bpr = 3 if rel > 1e-4 else 96
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "f.f32")
        data.tofile(path)
        out = api.run_ref_compress(path, list(shape), rel, os.path.join(tmp, "f.asz"),
                                   ["--block-pivot-ratio", str(bpr), "--predictor-plan", ",".join(plan)])
    want = re.search(r"^Predictor family selection: .*$", out, re.M).group(0)
    s = {"rel_eb": float(np.float32(rel)), "block_pivot_ratio": bpr, "eb_ratio": float(np.float32(0.76)),
         "pivot_eb_ratio": float(np.float32(0.05)), "slope_ratio": float(np.float32(0.54)), "t_precision": 32,
         "stat_eb_ratio": float(np.float32(0.2)), "h1_block_sampling_num": 256, "t_min": 1 / 16, "t_max": 6 / 16}
    line, rep = _select(ctx, data, s, plan)
    _say(line)
    _say("family selection time", rep.time_s, "s")
    assert line == want


CLI_CASES = [
    ("smooth_64x80x96_rel1e-3", (64, 80, 96), 1e-3, 3, lambda s: _grf(s, -5.0, 41)),
    ("turb_72x64x88_rel1e-4", (72, 64, 88), 1e-4, 96, lambda s: _grf(s, -11 / 3, 42, stretch=(4, 1, 2))),
]


@pytest.mark.parametrize("name,shape,rel,bpr,gen", CLI_CASES, ids=[c[0] for c in CLI_CASES])
def test_cli_automatic_pipeline_matches_reference(name, shape, rel, bpr, gen):
    """The drop-in s1_compress with neither --predictor-plan nor
    --predictor-family (plan search, family probe and compression all on the
    GPU, exact t scoring) prints the reference's search and family lines and
    writes the reference's .asz byte for byte."""
    import subprocess
    if not os.path.exists(api.REF_COMPRESS):
        pytest.skip("reference CLI not built (oracle/Makefile ref)")
    ours_bin = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "asz_amd", "bin",
                            "s1_compress")
    _say("case", name, shape, rel)
    data = gen(shape)
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "f.f32")
        data.tofile(path)
        ref_out = api.run_ref_compress(path, list(shape), rel, os.path.join(tmp, "ref.asz"),
                                       ["--block-pivot-ratio", str(bpr)])
        cmd = [ours_bin, "-f", "-i", path, "-3", *map(str, shape), "-M", "REL", f"{rel:.17g}", "-z",
               os.path.join(tmp, "ours.asz"), "--block-pivot-ratio", str(bpr), "--tsel", "exact"]
        ours_out = subprocess.run(cmd, capture_output=True, text=True, timeout=300, check=True).stdout
        keep = re.compile(r"^(best_layer_predictor=|Predictor family selection: |predictor_vec=|predictor_family="
                          r"|Default slope statistics|total_qb_num|t_int_vec)")
        assert [l for l in ours_out.splitlines() if keep.match(l)] == \
               [l for l in ref_out.splitlines() if keep.match(l)]
        with open(os.path.join(tmp, "ref.asz"), "rb") as fh:
            ref_asz = fh.read()
        with open(os.path.join(tmp, "ours.asz"), "rb") as fh:
            ours_asz = fh.read()
        assert ours_asz == ref_asz
