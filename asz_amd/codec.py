"""Python front end of the MI355X ASZ core (over the C ABI in include/asz_gpu.h).

Mirrors asz::param_c<float>'s interface for this path: a parameter block with
the reference's defaults (asz_param.hpp:147-183), a predictor plan of
"cubic<axis>_<order>" strings, compress() producing the pre-Zstd stream of
compress.tpp:243-291, decompress() consuming it (decompress.tpp:88-196), and
errors raised with the reference's messages.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import ASZ_TSEL_EXACT, ASZ_TSEL_GPU, ASZ_TSEL_GRAD, ASZ_TSEL_INJECT, Params, Plan, Stats

ORDERS = ("120", "021", "012")
TSEL = {"gpu": ASZ_TSEL_GPU, "exact": ASZ_TSEL_EXACT, "inject": ASZ_TSEL_INJECT, "grad": ASZ_TSEL_GRAD}


class AszError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def make_params(**kw) -> Params:
    p = Params()
    _lib.lib().asz_params_default(ctypes.byref(p))
    for k, v in kw.items():
        if k == "family":
            p.quintic = 1 if v == "quintic" else 0
            continue
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def make_plan(predictors) -> Plan:
    """["cubic0_012", ...] (or (axis, order) pairs) -> Plan."""
    pl = Plan()
    pl.layers = len(predictors)
    for r, pr in enumerate(predictors):
        if isinstance(pr, str):
            if len(pr) != 10 or not pr.startswith("cubic") or pr[5] not in "012" or pr[7:] not in ORDERS:
                raise ValueError(f"invalid predictor in plan: {pr}")
            pl.axis[r] = int(pr[5])
            pl.order[r] = ORDERS.index(pr[7:])
        else:
            pl.axis[r], pl.order[r] = int(pr[0]), int(pr[1])
    return pl


def plan_strings(axis, order):
    return [f"cubic{a}_{ORDERS[o]}" for a, o in zip(axis, order)]


@dataclass
class CompressResult:
    stream: bytes                       # pre-Zstd Huffman stream
    min_val: float
    max_val: float
    stats: Stats
    qb: np.ndarray | None = None        # pivot array then round arrays
    t: np.ndarray | None = None         # t_int per (block, round), block-major
    recon: np.ndarray | None = None
    counts: list = field(default_factory=list)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Context:
    """One HIP device and stream; calls are serialised on it."""

    def __init__(self, device: int = 0):
        self._L = _lib.lib()
        self._c = ctypes.c_void_p()
        rc = self._L.asz_ctx_create(device, ctypes.byref(self._c))
        if rc != 0:
            raise AszError(rc, "cannot create an ASZ context on HIP device %d" % device)

    def close(self):
        if self._c:
            self._L.asz_ctx_destroy(self._c)
            self._c = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise AszError(rc, self._L.asz_last_error(self._c).decode())

    def after_torch(self):
        """Order the context's next work after what torch has enqueued on its
        current stream (the inputs' producers), without a host wait."""
        import torch
        if torch.cuda.is_initialized():
            self._check(self._L.asz_stream_wait(self._c, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def set_kernel_events(self, on: bool = True):
        self._check(self._L.asz_set_option(self._c, _lib.ASZ_OPT_KERNEL_EVENTS, 1 if on else 0))

    def set_timed_round(self, r: int):
        """Time round r's kernel only (HIP events around that launch); r < 0: no events."""
        self._check(self._L.asz_set_option(self._c, _lib.ASZ_OPT_TIMED_ROUND, int(r)))

    def set_wide_codes(self, on: bool = True):
        """Route the Huffman encoder through its path for codes longer than 32 bits."""
        self._check(self._L.asz_set_option(self._c, _lib.ASZ_OPT_WIDE_CODES, 1 if on else 0))

    def set_coarse_rounds(self, n: int = -1):
        """The most leading rounds that run as block rounds (t selection and
        targets in one launch; -1: every round that qualifies, 0: none)."""
        self._check(self._L.asz_set_option(self._c, _lib.ASZ_OPT_COARSE_ROUNDS, int(n)))

    def set_graphs(self, on: bool = True):
        """Run multi-field compressions' device sequence as captured HIP graphs."""
        self._check(self._L.asz_set_option(self._c, _lib.ASZ_OPT_GRAPHS, 1 if on else 0))

    def autotune(self, field, plan: Plan, params: Params, reps: int = 3):
        """Time every candidate tile shape of each large round of `plan` on the
        CUDA tensor `field` (left untouched) and keep the fastest for this
        (shape, plan, block pivot ratio); returns the TuneReport."""
        rep = _lib.TuneReport()
        d = np.asarray(field.shape, np.uint64)
        self.after_torch()
        self._check(self._L.asz_autotune(self._c, ctypes.c_void_p(field.data_ptr()), _ptr(d), ctypes.byref(plan),
                                         ctypes.byref(params), int(reps), ctypes.byref(rep)))
        return rep

    def tune_lookup(self, shape, plan: Plan, params: Params):
        """The autotuned tile shapes for (shape, plan, block pivot ratio), or None."""
        t = (_lib.RoundTune * _lib.MAX_LAYERS)()
        d = np.asarray(shape, np.uint64)
        found = self._L.asz_tune_lookup(self._c, _ptr(d), ctypes.byref(plan), params.block_pivot_ratio,
                                        params.h1_block_sampling_num, t)
        return [(x.seg, x.xsh, x.rpw) for x in t[:plan.layers]] if found else None

    def save_tunes(self, path: str):
        self._check(self._L.asz_tune_save(self._c, path.encode()))

    def load_tunes(self, path: str):
        self._check(self._L.asz_tune_load(self._c, path.encode()))

    def set_stream_on_device(self, on: bool = True):
        """Leave compress_device's stream in HBM (stream_view() then copies it out)."""
        self._check(self._L.asz_set_option(self._c, _lib.ASZ_OPT_STREAM_ON_DEVICE, 1 if on else 0))
        self._stream_dev = on

    def stats(self) -> Stats:
        s = Stats()
        self._L.asz_get_stats(self._c, ctypes.byref(s))
        return s

    # ------------------------------------------------------------ device helpers
    def range(self, d_ptr: int, n: int):
        mn, mx = ctypes.c_float(), ctypes.c_float()
        self.after_torch()
        self._check(self._L.asz_range_f32(self._c, ctypes.c_void_p(d_ptr), n, ctypes.byref(mn), ctypes.byref(mx)))
        return mn.value, mx.value

    def default_slope(self, d_ptr: int, dims, params: Params, mn: float, mx: float):
        d = np.asarray(dims, np.uint64)
        s, inv = ctypes.c_float(), ctypes.c_float()
        fb, reason = ctypes.c_int(), ctypes.c_int()
        self.after_torch()
        self._check(self._L.asz_default_slope(self._c, ctypes.c_void_p(d_ptr), _ptr(d), ctypes.byref(params), mn, mx,
                                              ctypes.byref(s), ctypes.byref(inv), ctypes.byref(fb),
                                              ctypes.byref(reason)))
        return s.value, inv.value, bool(fb.value), ("none", "invalid_ratio", "invalid_difference")[reason.value]

    def generate_plan(self, d_ptr: int, dims, params: Params, mn: float, mx: float, max_layer_num=14,
                      full_predictor_search_layer_num=4, sample_stride=5):
        """Automatic plan (generate_config.tpp:443-678) for the field at d_ptr;
        params.default_slope must hold the field's default slope.  Returns the
        plan and the search report: per searched round (finest first) the
        winner's predictor name ("" when none had samples) and best_dist."""
        d = np.asarray(dims, np.uint64)
        plan, rep = Plan(), _lib.PlanReport()
        self.after_torch()
        self._check(self._L.asz_generate_plan(self._c, ctypes.c_void_p(d_ptr), _ptr(d), ctypes.byref(params), mn, mx,
                                              max_layer_num, full_predictor_search_layer_num, sample_stride,
                                              ctypes.byref(plan), ctypes.byref(rep)))
        lines = []
        for k in range(rep.searched):
            a = rep.best_axis[k]
            name = "" if a < 0 else f"cubic{a}_{ORDERS[rep.best_order[k]]}"
            lines.append((name, float(rep.best_dist[k])))
        return plan, {"searched": lines, "statistics_s": rep.statistics_s,
                      "sample_prediction_s": rep.sample_prediction_s,
                      "lossless_estimation_s": rep.lossless_estimation_s, "total_s": rep.total_s}

    def select_family(self, d_ptr: int, dims, plan: Plan, params: Params, mn: float, mx: float, suffix_layers=6,
                      sample_blocks=8, sample_point_limit=32768, sample_fraction_limit=0.02, fine_eb_threshold=1e-4,
                      coarse_probe_bpr=3, fine_probe_bpr=4):
        """Predictor family of the plan for the field at d_ptr
        (predictor_family.tpp:217-251); params.default_slope must hold the
        field's default slope.  Returns the report (a FamilyReport: the
        chosen family in .quintic and every value of the reference's
        "Predictor family selection" line)."""
        d = np.asarray(dims, np.uint64)
        o = _lib.FamilyOpts(suffix_layers, sample_blocks, sample_point_limit, float(np.float32(sample_fraction_limit)),
                            float(np.float32(fine_eb_threshold)), coarse_probe_bpr, fine_probe_bpr)
        rep = _lib.FamilyReport()
        self.after_torch()
        self._check(self._L.asz_select_family(self._c, ctypes.c_void_p(d_ptr), _ptr(d), ctypes.byref(plan),
                                              ctypes.byref(params), mn, mx, ctypes.byref(o), ctypes.byref(rep)))
        return rep

    def geometry(self, dims, plan: Plan, bpr: int):
        d = np.asarray(dims, np.uint64)
        ps, bs, bn = (np.zeros(3, np.uint64) for _ in range(3))
        self._check(self._L.asz_geometry(_ptr(d), ctypes.byref(plan), bpr, _ptr(ps), _ptr(bs), _ptr(bn)))
        return ps, bs, bn

    # ------------------------------------------------------------ compression
    def compress_device(self, d_ptr: int, dims, plan: Plan, params: Params, mn: float, mx: float, tsel="gpu",
                        t_inject=None, want_qb=False, want_t=False, copy_stream=True) -> CompressResult:
        """d_ptr: device float32 field, overwritten with the reconstruction.

        With copy_stream=False the result's stream is None and stream_view()
        returns the context's pinned buffer (valid until the next call)."""
        d = np.asarray(dims, np.uint64)
        N = int(np.prod(d))
        ti = None if t_inject is None else np.ascontiguousarray(t_inject, np.int64)
        qb = np.zeros(N, np.int32) if want_qb else None
        t = None
        if want_t:
            _, _, bn = self.geometry(dims, plan, params.block_pivot_ratio)
            t = np.zeros(int(np.prod(bn)) * plan.layers, np.int64)
        sp, sn = ctypes.c_void_p(), ctypes.c_uint64()
        self.after_torch()
        self._check(self._L.asz_compress(self._c, ctypes.c_void_p(d_ptr), _ptr(d), ctypes.byref(plan),
                                         ctypes.byref(params), mn, mx, TSEL[tsel], _ptr(ti), ctypes.byref(sp),
                                         ctypes.byref(sn), _ptr(qb), _ptr(t)))
        self._last_stream = (sp.value, sn.value)
        stream = None
        if copy_stream:
            stream = self.stream_view().tobytes() if getattr(self, "_stream_dev", False) else \
                ctypes.string_at(sp, sn.value)
        st = self.stats()
        return CompressResult(stream, mn, mx, st, qb, t, None, [int(st.qb_counts[k]) for k in range(plan.layers + 1)])

    def compress_field(self, d_ptr: int, dims, plan: Plan, params: Params, tsel="gpu",
                       copy_stream=True) -> CompressResult:
        """Range scan + compression of the device field at d_ptr in one call
        (asz_compress_field): the stream of range() + compress_device(), with
        no host wait between the scan and the compression."""
        d = np.asarray(dims, np.uint64)
        sp, sn = ctypes.c_void_p(), ctypes.c_uint64()
        mn, mx = ctypes.c_float(), ctypes.c_float()
        self.after_torch()
        self._check(self._L.asz_compress_field(self._c, ctypes.c_void_p(d_ptr), _ptr(d), ctypes.byref(plan),
                                               ctypes.byref(params), TSEL[tsel], ctypes.byref(sp), ctypes.byref(sn),
                                               ctypes.byref(mn), ctypes.byref(mx)))
        self._last_stream = (sp.value, sn.value)
        stream = None
        if copy_stream:
            stream = self.stream_view().tobytes() if getattr(self, "_stream_dev", False) else \
                ctypes.string_at(sp, sn.value)
        st = self.stats()
        return CompressResult(stream, mn.value, mx.value, st, None, None, None,
                              [int(st.qb_counts[k]) for k in range(plan.layers + 1)])

    def stream_device(self):
        """(device pointer, bytes) of the last compressed stream, left in this
        context's stream buffer by set_stream_on_device(True)."""
        if not getattr(self, "_stream_dev", False):
            raise AszError("the last stream is not in device memory (set_stream_on_device(True))")
        return self._last_stream

    def stream_view(self) -> np.ndarray:
        """The last compressed stream: the context's pinned buffer (no copy), or
        a host copy when the stream was left in device memory."""
        ptr, n = self._last_stream
        if getattr(self, "_stream_dev", False):
            out = np.empty(n, np.uint8)
            self._check(self._L.asz_copy_to_host(self._c, _ptr(out), ctypes.c_void_p(ptr), n))
            return out
        return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_uint8)), shape=(n,))

    def compress(self, field_host: np.ndarray, plan: Plan, params: Params, tsel="gpu", t_inject=None,
                 want_qb=False, want_t=False, want_recon=False) -> CompressResult:
        f = np.ascontiguousarray(field_host, np.float32)
        d = np.asarray(f.shape, np.uint64)
        _, _, bn = self.geometry(f.shape, plan, params.block_pivot_ratio)
        nt = int(np.prod(bn)) * plan.layers
        ti = None if t_inject is None else np.ascontiguousarray(t_inject, np.int64)
        qb = np.zeros(f.size, np.int32) if want_qb else None
        t = np.zeros(nt, np.int64) if want_t else None
        recon = np.zeros_like(f) if want_recon else None
        sp, sn = ctypes.c_void_p(), ctypes.c_uint64()
        self._check(self._L.asz_compress_host(self._c, _ptr(f), _ptr(d), ctypes.byref(plan), ctypes.byref(params),
                                              TSEL[tsel], _ptr(ti), ctypes.byref(sp), ctypes.byref(sn), _ptr(recon),
                                              _ptr(qb), _ptr(t)))
        stream = ctypes.string_at(sp, sn.value)
        st = self.stats()
        mn, mx = np.frombuffer(stream[:8], np.float32)
        return CompressResult(stream, float(mn), float(mx), st, qb, t, recon,
                              [int(st.qb_counts[k]) for k in range(plan.layers + 1)])

    def compress_batch(self, fields, plans, params: Params, tsel="grad", concurrency=6, on_device=False,
                       fields_per_chunk=0):
        """Compress many device fields (CUDA float32 tensors, overwritten with
        their reconstructions) in one call; plans: one Plan per field (or one
        for all).  Same-shape, same-plan fields go through the multi-field
        kernels in chunks of at most fields_per_chunk (0: automatic, 1: one
        field at a time).  Returns (streams, per-field BatchField info): bytes
        objects, or with on_device=True (device_pointer, nbytes) pairs the
        caller frees with free_device()."""
        n = len(fields)
        prev = getattr(self, "_stream_dev", False)
        self._check(self._L.asz_set_option(self._c, _lib.ASZ_OPT_BATCH_FIELDS, int(fields_per_chunk)))
        self._check(self._L.asz_set_option(self._c, _lib.ASZ_OPT_STREAM_ON_DEVICE, 1 if on_device else 0))
        if isinstance(plans, Plan):
            plans = [plans] * n
        ptrs = (ctypes.c_void_p * n)(*[f.data_ptr() for f in fields])
        dims = np.asarray([list(f.shape) for f in fields], np.uint64).reshape(-1)
        pl = (Plan * n)(*plans)
        outs = (ctypes.c_void_p * n)()
        nbytes = (ctypes.c_uint64 * n)()
        info = (_lib.BatchField * n)()
        try:
            self.after_torch()
            self._check(self._L.asz_compress_batch(self._c, n, ptrs, _ptr(dims), pl, ctypes.byref(params),
                                                   TSEL[tsel], concurrency, outs, nbytes, info))
        finally:
            self._L.asz_set_option(self._c, _lib.ASZ_OPT_STREAM_ON_DEVICE, 1 if prev else 0)
        streams = []
        for i in range(n):
            if on_device:
                streams.append((outs[i], int(nbytes[i])))
            else:
                streams.append(ctypes.string_at(outs[i], nbytes[i]))
                self._L.asz_free(outs[i])
        return streams, list(info)

    def free_device(self, ptr):
        self._check(self._L.asz_device_free(self._c, ctypes.c_void_p(ptr)))

    def device_bytes(self, ptr, n) -> bytes:
        out = np.empty(n, np.uint8)
        self._check(self._L.asz_copy_to_host(self._c, _ptr(out), ctypes.c_void_p(ptr), n))
        return out.tobytes()

    # ------------------------------------------------------------ decompression
    def decompress(self, stream: bytes, dims, params: Params, want_qb=False, want_t=False, nt=None):
        d = np.asarray(dims, np.uint64)
        N = int(np.prod(d))
        buf = np.frombuffer(stream, np.uint8)
        out = np.zeros(tuple(int(x) for x in dims), np.float32)
        if not want_qb and not want_t:
            self._check(self._L.asz_decompress_host(self._c, _ptr(buf), buf.size, _ptr(d), ctypes.byref(params),
                                                    _ptr(out)))
            return out
        import torch  # device buffer for the reconstruction
        dev = torch.empty(N, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        qb = np.zeros(N, np.int32) if want_qb else None
        t = np.zeros(nt, np.int64) if want_t else None
        self.after_torch()
        self._check(self._L.asz_decompress(self._c, _ptr(buf), buf.size, _ptr(d), ctypes.byref(params),
                                           ctypes.c_void_p(dev.data_ptr()), _ptr(qb), _ptr(t)))
        out[...] = dev.cpu().numpy().reshape(out.shape)
        return out, qb, t

    def decompress_device(self, d_stream: int, nbytes: int, dims, params: Params, d_out: int):
        """Decompresses a stream in device memory (d_stream, nbytes) into the
        device float32 buffer d_out.  A stream in this context's stream buffer
        (stream_view() after a compression with set_stream_on_device(True))
        is decoded where it lies."""
        d = np.asarray(dims, np.uint64)
        self.after_torch()
        self._check(self._L.asz_decompress_device(self._c, ctypes.c_void_p(d_stream), nbytes, _ptr(d),
                                                  ctypes.byref(params), ctypes.c_void_p(d_out)))


def family_line(rep) -> str:
    """The reference's "Predictor family selection: ..." line for a family
    report (predictor_family.tpp:237-247), long doubles printed by libc."""
    libc = ctypes.CDLL(None)

    def ld(name):
        v = ctypes.c_longdouble.from_buffer(rep, getattr(_lib.FamilyReport, name).offset)
        buf = ctypes.create_string_buffer(64)
        libc.snprintf(buf, 64, b"%.12Le", v)
        return buf.value.decode()

    head = (f"Predictor family selection: selector=h1, suffix_layers={rep.suffix_layers}, blocks={rep.blocks}, "
            f"sampled_block_points={rep.sampled_block_points}, sample_point_budget={rep.sample_point_budget}, "
            f"probe_block_pivot_ratio={rep.probe_block_pivot_ratio}")
    if rep.blocks == 0:
        return head + ", fallback=point_budget, family=cubic"
    return (head + f", points={rep.points}, cubic_mse={ld('cubic_mse')}, quintic_mse={ld('quintic_mse')}, "
            f"cubic_psnr={ld('cubic_psnr')}, quintic_psnr={ld('quintic_psnr')}, cubic_h1={ld('cubic_h1')}, "
            f"quintic_h1={ld('quintic_h1')}, six_score_gain={ld('score_gain')}, "
            f"family={'quintic' if rep.quintic else 'cubic'}")


def huffman_codebook(hist, offset: int, count: int, is_int64=False):
    """Host codebook + header bytes of one array (no GPU needed)."""
    L = _lib.lib()
    h = np.ascontiguousarray(hist, np.uint64)
    rng = h.size
    hdr = np.zeros(64 + 4 * rng * 8, np.uint8)
    hn = ctypes.c_uint64()
    codes = np.zeros(rng, np.uint64)
    lens = np.zeros(rng, np.uint8)
    bits = ctypes.c_uint64()
    rc = L.asz_huffman_codebook(_ptr(h), rng, offset, count, int(is_int64), _ptr(hdr), hdr.size, ctypes.byref(hn),
                                _ptr(codes), _ptr(lens), ctypes.byref(bits))
    if rc != 0:
        raise AszError(rc, "asz_huffman_codebook failed")
    return hdr[:hn.value].tobytes(), codes, lens, bits.value


def zstd_compress(payload: bytes, level: int = 3, threads: int = 1) -> bytes:
    """One Zstd frame around the pre-Zstd stream (compress.tpp:294-299: level 3,
    one thread).  threads > 1 compresses block-aligned slices concurrently and
    splices them into one frame (asz_zstd_compress_mt); 0 picks the thread count."""
    L = _lib.lib()
    buf = np.frombuffer(payload, np.uint8)
    out, n = ctypes.c_void_p(), ctypes.c_uint64()
    if threads == 1 and level == 3:
        rc = L.asz_zstd_compress(_ptr(buf), buf.size, ctypes.byref(out), ctypes.byref(n))
    else:
        rc = L.asz_zstd_compress_mt(_ptr(buf), buf.size, int(level), int(threads), ctypes.byref(out),
                                    ctypes.byref(n), None)
    if rc != 0:
        raise AszError(-2, "ZSTD_compress failed")
    try:
        return ctypes.string_at(out, n.value)
    finally:
        L.asz_free(out)


def zstd_decompress(frame: bytes) -> bytes:
    L = _lib.lib()
    buf = np.frombuffer(frame, np.uint8)
    out, n = ctypes.c_void_p(), ctypes.c_uint64()
    if L.asz_zstd_decompress(_ptr(buf), buf.size, ctypes.byref(out), ctypes.byref(n)) != 0:
        raise AszError(-2, "ZSTD_decompress error in decompress()")
    try:
        return ctypes.string_at(out, n.value)
    finally:
        L.asz_free(out)
