"""Fixture loading shared by the tests (reads tests/golden/fixtures only)."""
import json
import os

import numpy as np

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures")

_NUMERIC = {
    "--block-pivot-ratio": ("block_pivot_ratio", int),
    "--eb-ratio": ("eb_ratio", float),
    "--pivot-eb-ratio": ("pivot_eb_ratio", float),
    "--slope-ratio": ("slope_ratio", float),
    "--t-precision": ("t_precision", int),
    "--stat-eb-ratio": ("stat_eb_ratio", float),
    "--h1-block-sampling-num": ("h1_block_sampling_num", int),
    "--t-min": ("t_min", float),
    "--t-max": ("t_max", float),
}


def names():
    with open(os.path.join(FIX, "index.json")) as fh:
        return json.load(fh)


def load(name):
    with open(os.path.join(FIX, name + ".json")) as fh:
        info = json.load(fh)
    data = np.fromfile(os.path.join(FIX, name + ".f32"), np.float32).reshape(info["dims"])
    with open(os.path.join(FIX, name + ".asz"), "rb") as fh:
        asz = fh.read()
    return info, data, asz


def settings(info):
    """Numeric settings from the fixture's command line (as float32 where the CLI uses stof)."""
    out = {"rel_eb": float(np.float32(info["rel_eb"])), "block_pivot_ratio": 4, "eb_ratio": 0.76,
           "pivot_eb_ratio": 0.05, "slope_ratio": 0.54, "t_precision": 32, "stat_eb_ratio": 0.2,
           "h1_block_sampling_num": 256, "t_min": 1 / 16, "t_max": 6 / 16}
    a = info["args"]
    for i, tok in enumerate(a):
        if tok in _NUMERIC:
            key, typ = _NUMERIC[tok]
            out[key] = typ(a[i + 1])
    for k in ("eb_ratio", "pivot_eb_ratio", "slope_ratio", "stat_eb_ratio", "t_min", "t_max"):
        out[k] = float(np.float32(out[k]))
    return out
