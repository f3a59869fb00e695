"""Batched form: many fields of uneven shape in one call.

compress_fields() runs every field's pipeline (range scan, default slope,
rounds, Huffman) through asz_compress_batch.  Fields of one shape and plan
(time steps sharing a saved plan, same-shape fields) go through the
multi-field kernels in chunks, one launch per round for the whole chunk, and
worker contexts, each with its own HIP stream and host thread, run the chunks
concurrently so that the launch-latency-bound coarse rounds and the host
codebook work overlap on the device.  The config-4 driver (python -m
asz_amd.batch) compares it with the plain per-field form (one field after
another on one stream, discrete t scan) on 48 time steps of 100x500x500 plus
12 mixed-shape fields at REL 1e-3.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import codec, fields

# Plans the reference s1_compress chose (auto plan generation, cubic family,
# block pivot ratio 3, REL 1e-3) for step 0 of the time series (the saved plan
# every step reuses) and for the first field of each mixed shape.
C4_STEP_PLAN = ["cubic1_012", "cubic1_012", "cubic1_012", "cubic2_012", "cubic1_012", "cubic2_012", "cubic2_012",
                "cubic0_012", "cubic2_012", "cubic0_012", "cubic1_120", "cubic0_021", "cubic2_120", "cubic1_120"]
C4_SHAPE_PLANS = {
    (64, 256, 256): ["cubic0_012", "cubic2_012", "cubic0_012", "cubic2_012", "cubic0_012", "cubic2_012", "cubic1_012",
                     "cubic2_012", "cubic0_012", "cubic1_012", "cubic2_120", "cubic2_021", "cubic0_012", "cubic1_120"],
    (256, 384, 384): ["cubic0_012", "cubic2_012", "cubic2_012", "cubic0_012", "cubic0_012", "cubic2_012", "cubic0_012",
                      "cubic1_012", "cubic1_012", "cubic2_012", "cubic0_021", "cubic2_021", "cubic0_021", "cubic1_120"],
    (115, 69, 69): ["cubic2_012", "cubic0_012", "cubic2_012", "cubic1_012", "cubic2_012", "cubic1_012", "cubic0_012",
                    "cubic2_012", "cubic1_012", "cubic0_012", "cubic1_012", "cubic0_012", "cubic2_021", "cubic1_120"],
    (235, 449, 449): ["cubic0_012", "cubic1_012", "cubic1_012", "cubic1_012", "cubic2_012", "cubic2_012", "cubic0_012",
                      "cubic1_012", "cubic1_012", "cubic2_012", "cubic0_012", "cubic1_021", "cubic0_120", "cubic2_120"],
}


def c4_fields(steps=48, seed=4, device="cuda"):
    """48 shifted time steps of 100x500x500 and 12 mixed-shape fields, as CUDA tensors."""
    out = [s.contiguous() for s in fields.shifted_series((100, 500, 500), steps, seed=seed, device=device)]
    mixed = [(64, 256, 256)] * 4 + [(256, 384, 384)] * 4 + [(115, 69, 69)] * 2 + [(235, 449, 449)] * 2
    for k, shp in enumerate(mixed):
        g = fields.rescale(fields.grf(shp, -11.0 / 3.0, seed * 100 + k, device=device), 0.0, 1.0)
        out.append(g.to(torch.float32).contiguous())
    return out


def autotune(ctx: codec.Context, flds, plans, rel=1e-3, block_pivot_ratio=3, family="cubic", cache=None, reps=3):
    """Tile shapes for every distinct (shape, plan) of the batch: read from the
    cache file when it has them, otherwise autotuned on the first field of
    that shape (a copy; the fields are not modified) and written back.
    Returns the number of (shape, plan) keys tuned now."""
    import os
    params = codec.make_params(rel_eb=float(np.float32(rel)), block_pivot_ratio=block_pivot_ratio, family=family)
    if cache and os.path.exists(cache):
        ctx.load_tunes(cache)
    done, tuned = set(), 0
    for f, pl in zip(flds, plans):
        key = (tuple(f.shape), bytes(pl))
        if key in done:
            continue
        done.add(key)
        if ctx.tune_lookup(tuple(f.shape), pl, params) is not None:
            continue
        mn, mx = ctx.range(f.data_ptr(), f.numel())
        slope, *_ = ctx.default_slope(f.data_ptr(), tuple(f.shape), params, mn, mx)
        params.default_slope = slope
        ctx.autotune(f, pl, params, reps=reps)
        tuned += 1
    if cache and tuned:
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        ctx.save_tunes(cache)
    return tuned


def compress_fields(ctx: codec.Context, flds, plans, rel=1e-3, block_pivot_ratio=3, tsel="grad", concurrency=6,
                    on_device=False, fields_per_chunk=0, family="cubic"):
    """Batched form; returns (streams, per-field info, wall seconds).  With
    on_device the streams stay in HBM as (pointer, bytes) pairs;
    fields_per_chunk = 1 compresses one field at a time.  Tile shapes
    autotuned on ctx (autotune()) apply to the matching fields."""
    params = codec.make_params(rel_eb=float(np.float32(rel)), block_pivot_ratio=block_pivot_ratio, family=family)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    streams, info = ctx.compress_batch(flds, plans, params, tsel=tsel, concurrency=concurrency, on_device=on_device,
                                       fields_per_chunk=fields_per_chunk)
    return streams, info, time.perf_counter() - t0


def compress_plain(ctx: codec.Context, flds, plans, rel=1e-3, block_pivot_ratio=3, tsel="gpu"):
    """Plain per-field form: range scan, default slope and compression of one
    field after another on one stream; returns (streams, wall seconds)."""
    params = codec.make_params(rel_eb=float(np.float32(rel)), block_pivot_ratio=block_pivot_ratio, family="cubic")
    streams = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f, pl in zip(flds, plans):
        mn, mx = ctx.range(f.data_ptr(), f.numel())
        slope, *_ = ctx.default_slope(f.data_ptr(), tuple(f.shape), params, mn, mx)
        params.default_slope = slope
        res = ctx.compress_device(f.data_ptr(), tuple(f.shape), pl, params, mn, mx, tsel=tsel, copy_stream=True)
        streams.append(res.stream)
    torch.cuda.synchronize()
    return streams, time.perf_counter() - t0


def c4_plans(flds):
    return [codec.make_plan(C4_STEP_PLAN if tuple(f.shape) == (100, 500, 500) else C4_SHAPE_PLANS[tuple(f.shape)])
            for f in flds]


def config4_comparison(concurrency=6, reps=3, tune_cache=None):
    """Batched (gradient-tuned, concurrent) vs plain per-field GPU form on the
    60 config-4 fields at REL 1e-3: wall time, total compression ratio (host
    Zstd after the timed region) and max error over eb.  The plain form is
    the same C entry point with one worker (one field after another on one
    stream), the discrete t scan and the default tile shapes, so the
    comparison carries no Python overhead on either side; inputs and streams
    stay in HBM for both.  The batched forms run with the tile shapes the
    autotuner picked for each (shape, plan) (read from tune_cache when it has
    them; the tuning time is reported apart).  batched_per_field runs the
    batched configuration with one field per chunk (concurrency alone, no
    multi-field kernels)."""
    flds = c4_fields()
    plans = c4_plans(flds)
    n_bytes = sum(f.numel() * 4 for f in flds)
    ctx = codec.Context(0)
    plain_ctx = codec.Context(0)
    try:
        res = {}
        t0 = time.perf_counter()
        tuned_keys = autotune(ctx, flds, plans, cache=tune_cache)
        res["autotune"] = {"keys_tuned": tuned_keys, "seconds": time.perf_counter() - t0, "cache": tune_cache}
        for name, fn in (("plain", lambda w: compress_fields(plain_ctx, w, plans, tsel="gpu", concurrency=1,
                                                             on_device=True, fields_per_chunk=1)),
                         ("batched_tuned", lambda w: compress_fields(ctx, w, plans, tsel="grad",
                                                                     concurrency=concurrency, on_device=True)),
                         ("batched_plain_t", lambda w: compress_fields(ctx, w, plans, tsel="gpu",
                                                                       concurrency=concurrency, on_device=True)),
                         ("batched_per_field", lambda w: compress_fields(ctx, w, plans, tsel="grad",
                                                                         concurrency=concurrency, on_device=True,
                                                                         fields_per_chunk=1))):
            best = None
            for _ in range(reps):
                work = [f.clone() for f in flds]
                out = fn(work)
                streams, wall = out[0], out[-1]
                if best is None or wall < best[1]:
                    if best is not None:
                        for ptr, _ in best[0]:
                            ctx.free_device(ptr)
                    best = (streams, wall, work)
                else:
                    for ptr, _ in streams:
                        ctx.free_device(ptr)
            streams, wall, work = best
            z = 0
            for ptr, nb in streams:  # host Zstd after the timed region
                z += len(codec.zstd_compress(ctx.device_bytes(ptr, nb)))
                ctx.free_device(ptr)
            err = 0.0
            for f, w in zip(flds, work):
                mn, mx = float(f.min()), float(f.max())
                eb = float(np.float32(1e-3) * (np.float32(mx) - np.float32(mn)))
                err = max(err, float((w.double() - f.double()).abs().max()) / eb)
            res[name] = {"wall_s": wall, "GBps": n_bytes / wall / 1e9, "cr": n_bytes / z, "max_err_over_eb": err}
            del work
        res["speedup_tuned_batched_over_plain"] = res["plain"]["wall_s"] / res["batched_tuned"]["wall_s"]
        res["cr_tuned_over_plain"] = res["batched_tuned"]["cr"] / res["plain"]["cr"]
        res["fields"] = len(flds)
        res["bytes"] = n_bytes
        res["concurrency"] = concurrency
        return res
    finally:
        ctx.close()
        plain_ctx.close()


def run_batched(concurrency=6, reps=1, tsel="grad", tune_cache=None):
    """The batched form alone on the 60 config-4 fields (for profiling):
    returns the wall seconds of each repetition."""
    flds = c4_fields()
    plans = c4_plans(flds)
    with codec.Context(0) as ctx:
        autotune(ctx, flds, plans, cache=tune_cache)
        walls = []
        for _ in range(reps):
            work = [f.clone() for f in flds]
            streams, _, wall = compress_fields(ctx, work, plans, tsel=tsel, concurrency=concurrency, on_device=True)
            for ptr, _ in streams:
                ctx.free_device(ptr)
            walls.append(wall)
            del work
        return walls


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description="the 60-field batch: batched form vs plain per-field form")
    ap.add_argument("--concurrency", type=int, default=6)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--tune-cache", default=None)
    ap.add_argument("--batched-only", action="store_true", help="run only the batched gradient-tuned form")
    a = ap.parse_args()
    if a.batched_only:
        print(json.dumps({"wall_s": run_batched(a.concurrency, a.reps, tune_cache=a.tune_cache)}))
    else:
        print(json.dumps(config4_comparison(concurrency=a.concurrency, reps=a.reps, tune_cache=a.tune_cache),
                         indent=1))
