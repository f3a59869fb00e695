"""Write the golden fixtures under tests/golden/fixtures/.

Each fixture is a small float32 field (our own seeded generator), the .asz the
reference CLI (built from the reference's own sources by oracle/Makefile)
writes for it, and a JSON record with the command-line arguments, the plan and
default slope the reference printed, and the SHA-256 of the reconstruction the
reference s1_decompress writes.  Fields up to 64 KiB also keep that
reconstruction verbatim.

Run in a container that has /root/reference:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from asz_amd import fields  # noqa: E402
from oracle import api  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "fixtures")

ALL9 = "cubic0_120,cubic1_021,cubic2_012,cubic0_021,cubic1_012,cubic2_120,cubic0_012,cubic1_120,cubic2_021"


def grf(shape, exponent, seed, stretch=(1, 1, 1), lo=-80.0, hi=40.0):
    f = fields.grf(shape, exponent, seed, stretch=stretch)
    return fields.to_numpy(fields.rescale(f, lo, hi)).astype(np.float32)


def cases():
    rng = np.random.default_rng(7)
    out = []
    out.append(("tiny_7", grf((7, 7, 7), -11 / 3, 1), 1e-2, ["--block-pivot-ratio", "3"]))
    out.append(("small_9x13x17", grf((9, 13, 17), -11 / 3, 2), 1e-3, ["--block-pivot-ratio", "3"]))
    out.append(("mid_33x40x57_quintic", grf((33, 40, 57), -11 / 3, 3), 1e-3,
                ["--block-pivot-ratio", "3", "--predictor-family", "quintic"]))
    out.append(("mid_48x72x88_fine", grf((48, 72, 88), -11 / 3, 4, stretch=(4, 1, 2)), 1e-4,
                ["--block-pivot-ratio", "96"]))
    out.append(("cube64_allpred_cubic", grf((64, 64, 64), -11 / 3, 5, stretch=(4, 1, 2)), 1e-3,
                ["--block-pivot-ratio", "3", "--predictor-family", "cubic", "--predictor-plan", ALL9]))
    out.append(("cube64_allpred_quintic", grf((64, 64, 64), -3.0, 6), 1e-3,
                ["--block-pivot-ratio", "2", "--predictor-family", "quintic", "--predictor-plan", ALL9]))
    noise = rng.standard_normal((40, 50, 60)).astype(np.float32)
    out.append(("noise_40x50x60", noise, 1e-3, ["--block-pivot-ratio", "3"]))
    st = grf((50, 40, 30), -11 / 3, 8)
    i0, i1, i2 = np.meshgrid(np.arange(50), np.arange(40), np.arange(30), indexing="ij")
    st = (st + 36.0 * ((i0 - 25) + 2 * (i1 - 20) + 3 * (i2 - 15) > 0)).astype(np.float32)
    out.append(("step_50x40x30", st, 1e-2, ["--block-pivot-ratio", "3", "--predictor-family", "cubic"]))
    dn = grf((20, 21, 22), -11 / 3, 9, lo=-1.0, hi=1.0)
    flat = dn.reshape(-1)
    idx = rng.choice(flat.size, 400, replace=False)
    flat[idx[:200]] = np.float32(1e-40)
    flat[idx[200:300]] = np.float32(-3e-39)
    flat[idx[300:]] = 0.0
    out.append(("denormal_20x21x22", dn, 1e-3, ["--block-pivot-ratio", "3"]))
    sp = np.ones((16, 16, 16), np.float32)
    sp[5, 7, 9] = 2.0
    out.append(("spike_16", sp, 1e-2, ["--block-pivot-ratio", "3", "--predictor-family", "cubic"]))
    i0, i1, i2 = np.meshgrid(np.arange(24), np.arange(25), np.arange(26), indexing="ij")
    ramp = (0.5 * i0 - 0.25 * i1 + 0.125 * i2 + 0.01 * i0 * i1).astype(np.float32)
    out.append(("ramp_24x25x26", ramp, 1e-4, ["--block-pivot-ratio", "3", "--predictor-family", "cubic"]))
    out.append(("fine_32_rel1e-5", grf((32, 32, 32), -11 / 3, 10), 1e-5, ["--block-pivot-ratio", "96"]))
    out.append(("params_30x31x45", grf((30, 31, 45), -5 / 3, 11), 3e-3,
                ["--block-pivot-ratio", "2", "--eb-ratio", "0.8", "--pivot-eb-ratio", "0.1", "--slope-ratio", "0.5",
                 "--t-precision", "64", "--t-min", "0", "--t-max", "0.5", "--h1-block-sampling-num", "64",
                 "--stat-eb-ratio", "0.3", "--max-layer-num", "10", "--full-predictor-search-layer-num", "6"]))
    out.append(("rough_36x36x36_rel1e-1", grf((36, 36, 36), -5 / 3, 12), 1e-1, ["--block-pivot-ratio", "3"]))
    out.append(("thin_7x64x70", grf((7, 64, 70), -11 / 3, 13), 1e-3, ["--block-pivot-ratio", "3"]))
    return out


# flags that s1_decompress also accepts (s1_decompress.cpp:13-95)
DECOMP_FLAGS = {"--pivot-eb-ratio", "--eb-ratio", "--slope-ratio", "--block-pivot-ratio", "--max-layer-num",
                "--full-predictor-search-layer-num", "--t-precision", "--predictor-family"}


def decomp_args(extra):
    out = []
    i = 0
    while i < len(extra):
        if extra[i] in DECOMP_FLAGS:
            out += extra[i:i + 2]
        i += 2 if extra[i].startswith("--") and extra[i] != "--report-t-histogram" else 1
    return out


def main():
    os.makedirs(OUT, exist_ok=True)
    index = []
    with tempfile.TemporaryDirectory() as tmp:
        for name, field, rel, extra in cases():
            dims = list(field.shape)
            inp = os.path.join(OUT, name + ".f32")
            field.astype(np.float32).tofile(inp)
            asz = os.path.join(OUT, name + ".asz")
            rec = os.path.join(tmp, name + ".rec")
            cout = api.run_ref_compress(inp, dims, rel, asz, extra + ["--report-t-histogram"])
            dout = api.run_ref_decompress(asz, dims, rel, rec, decomp_args(extra))
            recon = np.fromfile(rec, np.float32)
            plan = re.search(r'^predictor_vec=\{(.*)\};$', cout, re.M).group(1)
            slope = re.search(r'^Default slope statistics: .*slope=([0-9.eE+-]+)$', cout, re.M).group(1)
            family = re.search(r'^predictor_family=(cubic|quintic);$', cout, re.M).group(1)
            rec_info = {
                "name": name, "dims": dims, "rel_eb": rel, "args": extra,
                "decompress_args": decomp_args(extra),
                "plan": re.findall(r'"([^"]+)"', plan), "family": family,
                "default_slope_printed": float(slope),
                "recon_sha256": hashlib.sha256(recon.tobytes()).hexdigest(),
                "asz_bytes": os.path.getsize(asz),
                "compress_stdout": [l for l in cout.splitlines() if re.match(
                    r"^(Default slope statistics|predictor_vec|predictor_family=|total_qb_num|t_int_vec|T histogram"
                    r"|T fallback|max_err|CR:|min_val|best_layer_predictor=|Predictor family selection: )", l)],
                "decompress_cr": float(re.search(r"^CR= ([0-9.eE+-]+)$", dout, re.M).group(1)),
            }
            if recon.nbytes <= 65536:
                recon.tofile(os.path.join(OUT, name + ".rec.f32"))
                rec_info["recon_file"] = name + ".rec.f32"
            with open(os.path.join(OUT, name + ".json"), "w") as fh:
                json.dump(rec_info, fh, indent=1)
            index.append(name)
            print(name, dims, rel, "plan", len(rec_info["plan"]), family, "bytes", rec_info["asz_bytes"])
    with open(os.path.join(OUT, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)


if __name__ == "__main__":
    main()
