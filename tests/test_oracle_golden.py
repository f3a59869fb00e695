"""Pins the oracle (the CPU restatement) against the reference CLI's own output.

Every fixture under tests/golden/fixtures was written by the reference
s1_compress / s1_decompress built from the reference's sources.  The oracle
must reproduce, bit for bit: the stream parse (all bytes consumed), the QB
arrays and the full pre-Zstd payload given the stream's plan and t, the
reconstruction (SHA-256 of the reference decompressor's output), the default
slope the reference printed, and the per-block t the reference selected.
"""
import hashlib

import numpy as np
import pytest

from asz_amd import codec
from oracle import api
from tests import golden_util as gu

NAMES = gu.names()


def _params(info, slope=1000.0):
    s = gu.settings(info)
    return api.Params(rel_eb=s["rel_eb"], pivot_eb_ratio=s["pivot_eb_ratio"], eb_ratio=s["eb_ratio"],
                      slope_ratio=s["slope_ratio"], block_pivot_ratio=s["block_pivot_ratio"],
                      t_precision=s["t_precision"], stat_eb_ratio=s["stat_eb_ratio"],
                      h1_block_sampling_num=s["h1_block_sampling_num"], default_slope=slope)


@pytest.mark.parametrize("name", NAMES)
def test_parse_consumes_whole_payload(name):
    info, data, asz = gu.load(name)
    payload = codec.zstd_decompress(asz)
    st = api.parse_payload(payload)
    assert st.nbytes == len(payload)
    assert int(st.counts.sum()) == data.size  # compress.tpp:230-231
    assert st.min_val == float(data.min()) and st.max_val == float(data.max())
    plan = [f"cubic{a}_{('120', '021', '012')[o]}" for a, o in zip(st.axis, st.order)]
    assert plan == info["plan"]
    assert st.quintic == (info["family"] == "quintic")


@pytest.mark.parametrize("name", NAMES)
def test_compress_with_stream_t_is_bit_exact(name):
    info, data, asz = gu.load(name)
    payload = codec.zstd_decompress(asz)
    st = api.parse_payload(payload)
    p = _params(info)
    t, counts, qb, recon, _ = api.compress(data, st.axis, st.order, st.quintic, p, t_inject=st.t,
                                           min_val=st.min_val, max_val=st.max_val)
    assert np.array_equal(counts, st.counts)
    assert np.array_equal(qb, st.qb)
    assert hashlib.sha256(recon.tobytes()).hexdigest() == info["recon_sha256"]
    rebuilt = api.build_payload(st.min_val, st.max_val, st.axis, st.order, st.quintic, t, counts, qb)
    assert rebuilt == payload


@pytest.mark.parametrize("name", NAMES)
def test_decompress_is_bit_exact(name):
    info, data, asz = gu.load(name)
    st = api.parse_payload(codec.zstd_decompress(asz))
    recon = api.decompress(st, data.shape, _params(info))
    assert hashlib.sha256(recon.tobytes()).hexdigest() == info["recon_sha256"]
    if "recon_file" in info:
        ref = np.fromfile(f"{gu.FIX}/{info['recon_file']}", np.float32).reshape(data.shape)
        assert np.array_equal(recon.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("name", NAMES)
def test_default_slope_and_t_selection(name):
    info, data, asz = gu.load(name)
    st = api.parse_payload(codec.zstd_decompress(asz))
    s = gu.settings(info)
    p = _params(info)
    slope, inv, fb, reason = api.default_slope(data, p, t_min=s["t_min"], t_max=s["t_max"])
    assert f"{slope:.6f}" == f"{info['default_slope_printed']:.6f}"
    line = [l for l in info["compress_stdout"] if l.startswith("Default slope statistics")][0]
    assert line == (f"Default slope statistics: max_invalid_ratio={inv:.6f}, fallback={'true' if fb else 'false'}, "
                    f"reason={reason}, slope={slope:.6f}")
    p.default_slope = slope
    t, counts, qb, recon, fallbacks = api.compress(data, st.axis, st.order, st.quintic, p,
                                                   min_val=st.min_val, max_val=st.max_val)
    assert np.array_equal(t, st.t)
    assert np.array_equal(qb, st.qb)
    assert f"T fallback count:{fallbacks}" in info["compress_stdout"]


def test_error_bound_with_float_rounding_slack():
    """|x - x'| <= eb + ulp_f32(x')/2 on every fixture (the reference's own reconstruction)."""
    for name in NAMES:
        info, data, asz = gu.load(name)
        st = api.parse_payload(codec.zstd_decompress(asz))
        recon = api.decompress(st, data.shape, _params(info))
        eb = float(api.abs_eb(gu.settings(info)["rel_eb"], st.min_val, st.max_val))
        err = np.abs(data.astype(np.float64) - recon.astype(np.float64))
        slack = np.spacing(np.abs(recon)).astype(np.float64) / 2
        assert np.all(err <= eb + slack + 1e-12 * eb), name
