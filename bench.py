#!/usr/bin/env python3
"""Benchmark of the MI355X ASZ compression core.

Main workload (one GPU, default "c2"): a 512^3 float32 field, anisotropic
|K|^-11/3 Gaussian random field (axis0 stretched 4x, axis2 2x) plus an oblique
planar step of 0.3 x range, compressed at REL 1e-3 with the cubic family and
the plan the reference s1_compress chose for this field.  One step = one full
compression of the field resident in HBM: range scan, pivots, every round (GPU
per-block t selection, prediction, quantization, reconstruction, QB
serialization) and Huffman coding (histogram, codebook, encode), with the
pre-Zstd stream left in HBM.  The same steps with the stream copied to host
memory are reported as compress_pcie_inclusive_GBps.  Zstd runs on the host
and is timed separately; the default slope (preprocessing) is computed once
before timing, as the reference excludes it from its compression time.

The same JSON line carries the other workloads (skipped with --no-extras):
  c3  the 1024^3 exp(GRF) density-like field at REL 1e-4 (block pivot ratio
      96): compress and decompress GB/s, the dominant kernel's roofline;
  c4  60 uneven fields (48 time steps of 100x500x500 sharing one saved plan
      and 12 mixed shapes) at REL 1e-3: the batched gradient-tuned form
      against the plain per-field GPU form (wall time, total CR, max error);
  c5  the 512^3 half-smooth / half-rough field at REL 1e-3 and 1e-4: the
      gradient-tuned t against the discrete t scan (CR, max error, t-selection
      kernel time).

Prints one JSON line: value = compression throughput in uncompressed GB/s.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Plans the reference s1_compress picked for these fields (auto plan generation,
# --predictor-family cubic, block pivot ratio by the launcher's policy).
PLAN_C2_REL_1E3 = ["cubic1_012", "cubic2_012", "cubic2_012", "cubic2_012", "cubic1_012", "cubic1_012", "cubic0_012",
                   "cubic2_012", "cubic2_012", "cubic0_012", "cubic1_021", "cubic0_012", "cubic2_012", "cubic0_120"]
PLAN_C3_REL_1E4 = ["cubic1_012", "cubic0_012", "cubic0_012", "cubic0_012", "cubic1_012", "cubic2_012", "cubic2_012",
                   "cubic1_012", "cubic1_012", "cubic2_012", "cubic0_120", "cubic0_120", "cubic1_120", "cubic2_120"]
WORKLOADS = {
    "c2": dict(shape=(512, 512, 512), rel=1e-3, plan=PLAN_C2_REL_1E3, field="config2", seed=2,
               desc="512^3 f32 anisotropic |K|^-11/3 GRF + oblique step, REL 1e-3"),
    "c3": dict(shape=(1024, 1024, 1024), rel=1e-4, plan=PLAN_C3_REL_1E4, field="config3", seed=3,
               desc="1024^3 f32 exp(|k|^-3 GRF) rescaled to [0, 1e3], REL 1e-4"),
}
HBM_PEAK_GBPS = 8000.0
STAGE_MIN_COUNT = 1 << 18  # rounds this large whose QB order runs along neither DIM2 nor the round axis are staged


def dom_kernel_name(plan, r, shape, counts):
    """Template name of the kernel that runs round r of a cubic plan (a large
    round whose order runs along neither DIM2 nor the round axis writes its
    QBs in "012" order and k_qb_order moves them into place; an axis-2 round
    with h = 1 on rows of a multiple of four elements runs as k_round_quad)."""
    axis, order = int(plan[r][5]), ("120", "021", "012").index(plan[r][7:])
    staged = order != 2 and counts[r + 1] >= STAGE_MIN_COUNT
    if axis == 2:
        h = 2 ** sum(1 for q in plan[r + 1:] if int(q[5]) == 2)
        kern = "k_round_quad" if h == 1 and shape[2] % 4 == 0 else "k_round_row"
        return f"{kern}<false, false, 2> + k_qb_order<true>" if staged else f"{kern}<false, false, {order}>"
    fastest = 2 if order == 2 else (0 if order == 0 else 1)
    wm = 0 if fastest == 2 else (1 if fastest == axis else 2)
    if wm == 2 and staged:
        return f"k_round_sweep<{axis}, false, false, 0> + k_qb_order<true>"
    return f"k_round_sweep<{axis}, false, false, {wm}>"


TRAFFIC_PROFILE = os.path.join("profiles", "hbm_traffic_512cube_rel1e-3.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on the 512^3 workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (each counter scaled by its
    reading on a plain copy kernel of known size); None if not measured."""
    path = os.path.join(ROOT, TRAFFIC_PROFILE)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            m = re.match(r"^\| `(.+?)` \| .*\| ([0-9]+) \|\s*$", line)
            if m and m.group(1) == kernel:
                return int(m.group(2))
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--tsel", default="gpu", choices=["gpu", "exact"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="only the main workload (no other-workload keys)")
    ap.add_argument("--json-out", default=None)
    return ap.parse_args()


def cpu_baseline(field_np, plan, rel, bpr):
    """Reference s1_compress / s1_decompress (built from the reference's sources
    into oracle/_ref) on the same field, one core, pinned with taskset."""
    from oracle import api
    if not api.ref_available():
        return None
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        inp = os.path.join(tmp, "field.f32")
        field_np.tofile(inp)
        asz = os.path.join(tmp, "field.asz")
        rec = os.path.join(tmp, "field.rec")
        dims = [str(x) for x in field_np.shape]
        common = ["-f", "-3", *dims, "-M", "REL", f"{rel:.17g}", "--block-pivot-ratio", str(bpr),
                  "--predictor-family", "cubic"]
        pin = ["taskset", "-c", "0"] if subprocess.run(["which", "taskset"], capture_output=True).returncode == 0 else []
        t0 = time.time()
        out = subprocess.run(pin + [api.REF_COMPRESS, "-i", inp, "-z", asz, *common, "--predictor-plan", ",".join(plan)],
                             capture_output=True, text=True, timeout=600, check=True).stdout
        wall_c = time.time() - t0
        dout = subprocess.run(pin + [api.REF_DECOMPRESS, "-z", asz, "-o", rec, *common], capture_output=True,
                              text=True, timeout=600, check=True).stdout
        ctime = float(re.search(r"^Compression time: ([0-9.eE+-]+)", out, re.M).group(1))
        dtime = float(re.search(r"^Decompression time: ([0-9.eE+-]+)", dout, re.M).group(1))
        cr = float(re.search(r"^CR= ([0-9.eE+-]+)", dout, re.M).group(1))
        with open(asz, "rb") as fh:
            ref_frame = fh.read()
    nbytes = field_np.nbytes
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            cpu = re.search(r"^model name\s*:\s*(.*)$", fh.read(), re.M).group(1).strip()
    except Exception:
        pass
    return {"value": nbytes / ctime / 1e9, "unit": "GB/s", "cores": 1, "kind": "reference",
            "sample": f"whole {'x'.join(map(str, field_np.shape))} field, same plan/REL/family; reference "
                      f"'Compression time' (range + core, no preprocessing) {ctime:.3f} s, wall {wall_c:.2f} s",
            "decompress_GBps": nbytes / dtime / 1e9, "cr": cr, "cpu_model": cpu, "_frame": ref_frame}


def measure(W, steps, warmup, tsel="gpu", traffic_lookup=False):
    """Compression and decompression of workload W on cuda:0.  Returns
    (result dict, field tensor, context, params, abs_eb, stream)."""
    import numpy as np
    import torch

    from asz_amd import codec, fields

    rel = W["rel"]
    shape = tuple(W["shape"])
    N = int(np.prod(shape))
    nbytes = N * 4
    bpr = 3 if rel > 1e-4 else 96
    plan = W["plan"]
    gen = fields.config2_field if W["field"] == "config2" else fields.config3_field
    field = gen(shape, seed=W["seed"], device="cuda").contiguous()
    torch.cuda.synchronize()

    ctx = codec.Context(0)
    params = codec.make_params(rel_eb=float(np.float32(rel)), block_pivot_ratio=bpr, family="cubic")
    pl = codec.make_plan(plan)
    mn, mx = ctx.range(field.data_ptr(), N)
    slope, *_ = ctx.default_slope(field.data_ptr(), shape, params, mn, mx)
    params.default_slope = slope

    # inputs resident in HBM: one copy per step (compression overwrites its input with the reconstruction)
    works = [field.clone() for _ in range(max(warmup + steps, 2))]
    torch.cuda.synchronize()

    def step(w):
        if tsel == "gpu":
            # range scan + compression in one call (asz_compress_field): the
            # range and the bounds derived from it stay on the device
            return ctx.compress_field(w.data_ptr(), shape, pl, params, tsel=tsel, copy_stream=False)
        m0, m1 = ctx.range(w.data_ptr(), N)
        return ctx.compress_device(w.data_ptr(), shape, pl, params, m0, m1, tsel=tsel, copy_stream=False)

    def refill(k):
        for i in range(k):
            works[i].copy_(field)
        torch.cuda.synchronize()

    ctx.set_stream_on_device(True)  # the stream stays in HBM: inputs and outputs resident
    for k in range(warmup):
        step(works[k])
    # Untimed, after the warmup: one step with an event pair around every
    # launch; the dominant kernel is the compress launch with the longest
    # device time (a round launch here; the t selection, fold and encoder
    # launches are shorter, as the per-kernel breakdown below shows).
    probe = field.clone()
    ctx.set_kernel_events(True)
    res = step(probe)
    ctx.set_kernel_events(False)
    del probe
    counts = res.counts
    dom = max(range(len(plan)), key=lambda r: res.stats.round_ms[r])
    # timed region: events only around the dominant round's kernel
    ctx.set_timed_round(dom)
    dom_ms = 0.0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        res = step(works[warmup + k])
        dom_ms += res.stats.round_ms[dom]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.set_timed_round(-1)
    ms_per_step = (t1 - t0) * 1e3 / steps
    value = nbytes / (ms_per_step / 1e3) / 1e9
    dom_ms /= steps
    st = res.stats
    stream = ctx.stream_view().tobytes()  # copied out of HBM after the timed region
    S_h = len(stream)

    # per-kernel breakdown (untimed: an event pair around every launch)
    K = 2
    refill(K)
    ctx.set_kernel_events(True)
    acc = {"range": 0.0, "pivot": 0.0, "hist": 0.0, "encode": 0.0}
    round_ms = [0.0] * len(plan)
    tsel_ms = [0.0] * len(plan)
    block_rounds = 0
    for k in range(K):
        bst = step(works[k]).stats
        acc["range"] += bst.range_ms / K
        acc["pivot"] += bst.pivot_ms / K
        acc["hist"] += bst.hist_ms / K
        acc["encode"] += bst.encode_ms / K
        block_rounds = int(bst.coarse_rounds)
        for r in range(len(plan)):
            round_ms[r] += bst.round_ms[r] / K
            tsel_ms[r] += bst.tsel_round_ms[r] / K
    ctx.set_kernel_events(False)

    # the same steps with the stream copied to page-locked host memory (PCIe-inclusive rate)
    ctx.set_stream_on_device(False)
    pcie_steps = min(steps, len(works))
    refill(pcie_steps)
    p0 = time.perf_counter()
    for k in range(pcie_steps):
        step(works[k])
    torch.cuda.synchronize()
    pcie_ms = (time.perf_counter() - p0) * 1e3 / pcie_steps
    pcie_same_stream = ctx.stream_view().tobytes() == stream

    # A plain copy kernel of known size (reads and writes 4N bytes): the
    # reference point that scales rocprof's FETCH_SIZE / WRITE_SIZE counters.
    for _ in range(2):
        torch.mul(field, 1.0, out=works[0])
    torch.cuda.synchronize()
    del works

    # Zstd on the host, timed separately
    z0 = time.perf_counter()
    frame = codec.zstd_compress(stream)
    zstd_ms = (time.perf_counter() - z0) * 1e3
    cr = nbytes / len(frame)

    # decompression throughput (GPU Huffman decode + reconstruction): from the
    # stream in a device buffer (only its header bytes go to the host), and
    # PCIe-inclusive from page-locked host memory
    dstream = torch.frombuffer(bytearray(stream), dtype=torch.uint8).cuda()
    dev_ptr, dev_n = dstream.data_ptr(), dstream.numel()
    out = torch.empty(N, dtype=torch.float32, device="cuda")
    pinned = torch.empty(S_h, dtype=torch.uint8).pin_memory()  # the stream in page-locked host memory
    pinned.numpy()[:] = np.frombuffer(stream, np.uint8)
    dims_arr = np.asarray(shape, np.uint64)
    torch.cuda.synchronize()
    dec_ms = []

    def dec_once():
        ctx._check(ctx._L.asz_decompress(ctx._c, codec.ctypes.c_void_p(pinned.data_ptr()), S_h,
                                         dims_arr.ctypes.data_as(codec.ctypes.c_void_p), codec.ctypes.byref(params),
                                         codec.ctypes.c_void_p(out.data_ptr()), None, None))

    for k in range(warmup + steps):
        d0 = time.perf_counter()
        dec_once()
        d1 = time.perf_counter()
        if k >= warmup:
            dec_ms.append((d1 - d0) * 1e3)
    dec_pcie_ms_step = sum(dec_ms) / len(dec_ms)
    dec_ms = []
    for k in range(warmup + steps):
        d0 = time.perf_counter()
        ctx.decompress_device(dev_ptr, dev_n, shape, params, out.data_ptr())
        torch.cuda.synchronize()
        d1 = time.perf_counter()
        if k >= warmup:
            dec_ms.append((d1 - d0) * 1e3)
    dec_round = [0.0] * len(plan)
    dec_huff = 0.0
    ctx.set_kernel_events(True)
    for k in range(K):
        dec_once()
        ds = ctx.stats()
        dec_huff += ds.decode_ms / K
        for r in range(len(plan)):
            dec_round[r] += ds.round_ms[r] / K
    ctx.set_kernel_events(False)
    dec_ms_step = sum(dec_ms) / len(dec_ms)
    recon = out.view(shape)
    max_err = 0.0
    sq = 0.0
    for s0 in range(0, shape[0], 64):  # slab by slab (bounded temporaries at 1024^3)
        e = (recon[s0:s0 + 64].double() - field[s0:s0 + 64].double()).abs()
        max_err = max(max_err, float(e.max()))
        sq += float((e * e).sum())
    mse = sq / N
    rng = float(mx) - float(mn)
    psnr = 20 * np.log10(rng) - 10 * np.log10(mse) if mse > 0 else float("inf")
    abs_eb = float(np.float32(rel) * (np.float32(mx) - np.float32(mn)))
    del out, recon, pinned, dstream

    # roofline of the dominant kernel, timed with HIP events on the library's
    # stream inside the timed region; algorithmic bytes per target = 16
    # (original read, coarse read, reconstruction write, QB write)
    T_dom = counts[dom + 1]
    dom_bytes = 16 * T_dom
    dom_gbps = dom_bytes / (dom_ms / 1e3) / 1e9
    kname = dom_kernel_name(plan, dom, shape, counts)
    P = counts[0]
    gpu_ms = acc["range"] + acc["pivot"] + sum(round_ms) + sum(tsel_ms) + acc["hist"] + acc["encode"]
    B_c = 28 * N - 4 * P + S_h
    dec_gpu_ms = dec_huff + sum(dec_round)
    B_d = 16 * N - 4 * P + S_h
    launches = sorted([(f"round {r}", round_ms[r]) for r in range(len(plan))] +
                      [(f"tsel {r}", tsel_ms[r]) for r in range(len(plan)) if tsel_ms[r] > 0] +
                      [("range", acc["range"])], key=lambda x: -x[1])[:4]
    result = {
        "value": round(value, 3),
        "ms_per_step": round(ms_per_step, 4),
        "config": {"workload": W["desc"] + ", cubic, fixed reference plan, block_pivot_ratio "
                               f"{bpr}, GPU t selection; one step = range + compress to the pre-Zstd stream",
                   "shape": list(shape), "rel_eb": rel, "plan": plan, "family": "cubic", "block_pivot_ratio": bpr,
                   "tsel": tsel},
        "roofline": {"bound": "hbm", "achieved": round(dom_gbps, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(dom_gbps / HBM_PEAK_GBPS, 4),
                     "traffic": measured_traffic(kname) if traffic_lookup else None,
                     "traffic_source": (TRAFFIC_PROFILE + " (rocprofv3 FETCH_SIZE and WRITE_SIZE passes of bench.py "
                                        "at the commit of these kernels; not measured in this run)")
                     if traffic_lookup else None,
                     "kernel": f"{kname} (round {dom}, axis {plan[dom][5]}, order {plan[dom][7:]}; the longest "
                               "compress launch)",
                     "algorithmic_bytes": dom_bytes, "avg_ms": round(dom_ms, 5)},
        "compress_pcie_inclusive_GBps": round(nbytes / (pcie_ms / 1e3) / 1e9, 3),
        "pcie_leg_same_stream": pcie_same_stream,
        "decompress_GBps": round(nbytes / (dec_ms_step / 1e3) / 1e9, 3),
        "decompress_ms_per_step": round(dec_ms_step, 4),
        "decompress_pcie_inclusive_GBps": round(nbytes / (dec_pcie_ms_step / 1e3) / 1e9, 3),
        "pipeline": {
            "compress_gpu_kernel_ms": round(gpu_ms, 4),
            "compress_algorithmic_bytes": B_c,
            "compress_pipeline_roofline_frac": round(B_c / (gpu_ms / 1e3) / 1e9 / HBM_PEAK_GBPS, 4),
            "compress_wall_roofline_frac": round(B_c / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBPS, 4),
            "decompress_gpu_kernel_ms": round(dec_gpu_ms, 4),
            "decompress_algorithmic_bytes": B_d,
            "decompress_pipeline_roofline_frac": round(B_d / (dec_gpu_ms / 1e3) / 1e9 / HBM_PEAK_GBPS, 4),
            "decompress_wall_roofline_frac": round(B_d / (dec_ms_step / 1e3) / 1e9 / HBM_PEAK_GBPS, 4),
            "longest_launches_ms": [[n, round(v, 4)] for n, v in launches],
            "block_rounds": block_rounds,
            "range_ms": round(acc["range"], 4),
            "round_ms": [round(x, 4) for x in round_ms],
            "tsel_ms": [round(x, 4) for x in tsel_ms],
            "pivot_ms": round(acc["pivot"], 4), "hist_fold_ms": round(acc["hist"], 4),
            "encode_ms": round(acc["encode"], 4), "huffman_host_ms": round(st.huffman_host_ms, 4),
            "decode_ms": round(dec_huff, 4), "dec_round_ms": [round(x, 4) for x in dec_round],
            "zstd_host_ms": round(zstd_ms, 3),
        },
        "cr": round(cr, 4), "huffman_stream_bytes": S_h, "asz_bytes": len(frame),
        "max_err_over_eb": round(max_err / abs_eb, 7), "psnr": round(psnr, 4),
        "t_fallbacks": int(st.t_fallbacks),
    }
    return result, field, ctx, params, abs_eb


def c3_line():
    """The 1024^3 workload: compress / decompress rates and its dominant kernel."""
    import torch
    r, field, ctx, _, _ = measure(WORKLOADS["c3"], steps=3, warmup=1)
    ctx.close()
    del field
    torch.cuda.empty_cache()
    keep = ("value", "ms_per_step", "roofline", "decompress_GBps", "decompress_ms_per_step", "cr", "max_err_over_eb",
            "psnr")
    out = {k: r[k] for k in keep}
    out["unit"] = "GB/s"
    out["workload"] = r["config"]["workload"]
    p = r["pipeline"]
    for k in ("compress_pipeline_roofline_frac", "compress_wall_roofline_frac", "decompress_pipeline_roofline_frac",
              "longest_launches_ms"):
        out[k] = p[k]
    return out


def c4_line():
    """The 60-field batch: batched gradient-tuned form vs plain per-field GPU form."""
    import torch
    from asz_amd import batch
    # tile shapes autotuned once per (shape, plan) into a cache file, then reused
    r = batch.config4_comparison(reps=3, tune_cache=os.path.join(tempfile.gettempdir(), "asz_amd_tunes.txt"))
    torch.cuda.empty_cache()
    keep = {"plain": r["plain"], "batched_tuned": r["batched_tuned"], "batched_plain_t": r["batched_plain_t"],
            "batched_per_field": r["batched_per_field"], "autotune": r["autotune"]}
    return {"workload": "48 time steps of 100x500x500 (one saved plan) + 12 mixed shapes, REL 1e-3, "
                        f"{r['fields']} fields, {r['bytes'] / 1e9:.2f} GB; best of 3 runs each",
            **keep,
            "speedup_tuned_batched_over_plain": r["speedup_tuned_batched_over_plain"],
            "cr_tuned_over_plain": r["cr_tuned_over_plain"], "concurrency": r["concurrency"]}


def c5_line():
    """The half-smooth / half-rough 512^3 field: gradient-tuned t vs the discrete t scan."""
    import torch
    from asz_amd import tune
    r = tune.config5_comparison(rels=(1e-3, 1e-4))
    torch.cuda.empty_cache()
    return {"workload": "512^3 f32, half |k|^-11/3 (smooth), half |k|^-5/3 (rough) along axis 0",
            "rel_1e-3": r["0.001"], "rel_1e-4": r["0.0001"]}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {a.gpus} is not supported (every run here has a single MI355X)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from asz_amd import codec

    torch.cuda.set_device(0)
    W = WORKLOADS[a.workload]
    r, field, ctx, params, abs_eb = measure(W, a.steps, a.warmup, a.tsel, traffic_lookup=a.workload == "c2")
    line = {
        "metric": "compress/decompress GB/s + % HBM roofline; CR at REL 1e-3, tuned vs plain",
        "value": r["value"],
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded GRF generated on the GPU)",
    }
    line.update({k: v for k, v in r.items() if k not in line})
    if not a.no_cpu_baseline:
        cfg = r["config"]
        cb = cpu_baseline(field.cpu().numpy(), cfg["plan"], cfg["rel_eb"], cfg["block_pivot_ratio"])
        if cb is not None:
            ref_frame = cb.pop("_frame")
            line["cpu_baseline"] = cb
            line["cr_reference"] = cb.pop("cr")
            line["cr_vs_reference"] = round(r["cr"] / line["cr_reference"], 5)
            # cross-decode: the reference stream through the GPU decoder
            ref_out = ctx.decompress(codec.zstd_decompress(ref_frame), tuple(cfg["shape"]), params)
            rerr = float(np.max(np.abs(ref_out.astype(np.float64) - field.cpu().numpy().astype(np.float64))))
            line["reference_stream_gpu_decoded_max_err_over_eb"] = round(rerr / abs_eb, 7)
    else:
        line["cpu_baseline"] = None
    ctx.close()
    del field
    torch.cuda.empty_cache()
    if a.workload == "c2" and not a.no_extras:
        for key, fn in (("c4", c4_line), ("c5", c5_line), ("c3", c3_line)):
            try:
                line[key] = fn()
            except Exception as e:  # reported, not fatal: the main line stands on its own
                line[key] = {"error": f"{type(e).__name__}: {e}"}
            torch.cuda.empty_cache()
    txt = json.dumps(line)
    print(txt, flush=True)
    if a.json_out:
        with open(a.json_out, "w") as fh:
            fh.write(txt + "\n")


if __name__ == "__main__":
    main()
