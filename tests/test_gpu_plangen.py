"""Automatic plan generation on the GPU (asz_generate_plan) against the
reference's own search, param_c::generate_config (generate_config.tpp:443-678).

Bar: the same plan, and every "best_layer_predictor=..., best_dist=..." line
the reference s1_compress prints (one per searched round, best_dist printed
with %f, so the float psnr - slope * br of each winner agrees to the printed
digits), on
* the golden fixtures whose command line has no --predictor-plan (the
  reference's lines are recorded in the fixture), and
* larger generated fields, compared with the reference CLI built from its
  own sources (oracle/_ref/s1_compress) run on the same file.
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

NAMES = [n for n in gu.names() if "--predictor-plan" not in gu.load(n)[0]["args"]]


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


def _search_args(args):
    out = {"max_layer_num": 14, "full_predictor_search_layer_num": 4, "sample_stride": 5}
    keys = {"--max-layer-num": "max_layer_num", "--full-predictor-search-layer-num": "full_predictor_search_layer_num",
            "--generate-config-sample-stride": "sample_stride"}
    for i, tok in enumerate(args):
        if tok in keys:
            out[keys[tok]] = int(args[i + 1])
    return out


def _generate(ctx, data, settings, search):
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
    plan, rep = ctx.generate_plan(t.data_ptr(), data.shape, p, mn, mx, **search)
    lines = [f"best_layer_predictor={n}, best_dist={d:.6f}" for n, d in rep["searched"]]
    return codec.plan_strings(list(plan.axis[:plan.layers]), list(plan.order[:plan.layers])), lines, rep


@pytest.mark.parametrize("name", NAMES)
def test_plan_matches_reference_search(ctx, name):
    info, data, _ = gu.load(name)
    _say("case", name, data.shape)
    plan, lines, rep = _generate(ctx, data, gu.settings(info), _search_args(info["args"]))
    want = [l for l in info["compress_stdout"] if l.startswith("best_layer_predictor=")]
    assert lines == want
    assert plan == info["plan"]


def _grf(shape, exponent, seed, stretch=(1, 1, 1)):
    f = fields.rescale(fields.grf(shape, exponent, seed, stretch=stretch), -80.0, 40.0)
    return fields.to_numpy(f).astype(np.float32)


GENERATED = [
    ("grf_64x96x128", (64, 96, 128), 1e-3, 3, lambda s: _grf(s, -11 / 3, 21, stretch=(4, 1, 2))),
    ("grf_80x72x100_rel1e-4", (80, 72, 100), 1e-4, 96, lambda s: _grf(s, -3.0, 22)),
    ("rough_50x120x90_rel1e-2", (50, 120, 90), 1e-2, 3, lambda s: _grf(s, -5 / 3, 23)),
    ("config1_100x500x500", (100, 500, 500), 1e-3, 3,
     lambda s: fields.to_numpy(fields.config1_field(s)).astype(np.float32)),
]


@pytest.mark.parametrize("name,shape,rel,bpr,gen", GENERATED, ids=[g[0] for g in GENERATED])
def test_plan_matches_reference_binary(ctx, name, shape, rel, bpr, gen):
    if not os.path.exists(api.REF_COMPRESS):
        pytest.skip("reference CLI not built (oracle/Makefile ref)")
    _say("case", name, shape, rel)
    data = gen(shape)
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "f.f32")
        data.tofile(path)
        out = api.run_ref_compress(path, list(shape), rel, os.path.join(tmp, "f.asz"),
                                   ["--block-pivot-ratio", str(bpr), "--predictor-family", "cubic"])
    want_lines = re.findall(r"^best_layer_predictor=.*$", out, re.M)
    want_plan = re.findall(r'"([^"]+)"', re.search(r"^predictor_vec=\{(.*)\};$", out, re.M).group(1))
    s = {"rel_eb": float(np.float32(rel)), "block_pivot_ratio": bpr, "eb_ratio": float(np.float32(0.76)),
         "pivot_eb_ratio": float(np.float32(0.05)), "slope_ratio": float(np.float32(0.54)), "t_precision": 32,
         "stat_eb_ratio": float(np.float32(0.2)), "h1_block_sampling_num": 256, "t_min": 1 / 16, "t_max": 6 / 16}
    plan, lines, rep = _generate(ctx, data, s, _search_args([]))
    _say("generate config time", rep["total_s"], "s")
    assert lines == want_lines
    assert plan == want_plan
