/*
 * asz_gpu.h — C ABI of the MI355X ASZ compression core (libasz_gpu.so).
 *
 * The reference has no library ABI: its core is header-only C++ behind two
 * executables (c_codes/s1_compress.cpp, c_codes/s1_decompress.cpp).  Each
 * entry point below replaces one member of asz::param_c<float> or one step of
 * those mains; the reference location is given with it.  Plain pointers and
 * sizes only.  Every call returns ASZ_OK (0) or a negative error code;
 * asz_last_error(ctx) then holds the message (the text the reference's
 * exceptions carry where one exists).  Pointers named d_* are HIP device
 * pointers on the context's device; all others are host pointers.
 *
 * One context per host thread; calls on one context are serialised on its
 * HIP stream.  Contexts are not shared between threads.
 */
#ifndef ASZ_GPU_H
#define ASZ_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ASZ_OK 0
#define ASZ_ERR_ARG (-1)     /* invalid argument (std::invalid_argument in the reference) */
#define ASZ_ERR_DATA (-2)    /* data / stream error (std::runtime_error in the reference)  */
#define ASZ_ERR_HIP (-3)     /* HIP runtime failure                                         */
#define ASZ_ERR_NOMEM (-4)   /* device or host allocation failure                           */

#define ASZ_MAX_LAYERS 63

typedef struct asz_ctx asz_ctx;

/* The numeric knobs of param_c<float> that the core reads (asz_param.hpp:147-183).
 * asz_params_default fills the reference defaults (block_pivot_ratio 4). */
typedef struct asz_params {
    float rel_eb;                   /* -M REL x                          (s1_compress.cpp:191-196) */
    float pivot_eb_ratio;           /* --pivot-eb-ratio      default 0.05 */
    float eb_ratio;                 /* --eb-ratio            default 0.76 */
    float slope_ratio;              /* --slope-ratio         default 0.54 */
    uint64_t block_pivot_ratio;     /* --block-pivot-ratio   default 4    */
    int32_t t_precision;            /* --t-precision         default 32   */
    float stat_eb_ratio;            /* --stat-eb-ratio       default 0.2  */
    uint64_t h1_block_sampling_num; /* --h1-block-sampling-num default 256 */
    float t_min, t_max;             /* --t-min / --t-max     default 1/16, 6/16 */
    float default_slope_eb_span;    /* --default-slope-eb-span            default 1.5  */
    float default_slope_invalid_ratio_threshold; /* default 1 */
    float default_slope_fallback;   /* --default-slope-fallback           default 1000 */
    float default_slope;            /* lambda of the RD proxy, from asz_default_slope */
    int32_t quintic;                /* predictor family: 0 cubic, 1 quintic (force_quintic) */
} asz_params;

/* A predictor plan: one (axis, QB order) pair per round, coarsest first.
 * order 0 = "120", 1 = "021", 2 = "012" (asz_param.hpp:122-129). */
typedef struct asz_plan {
    int32_t layers;
    uint8_t axis[ASZ_MAX_LAYERS];
    uint8_t order[ASZ_MAX_LAYERS];
} asz_plan;

/* Per-block t source for asz_compress. */
#define ASZ_TSEL_GPU 0      /* GPU H1-PSNR selection (compress.tpp:90-144 on the GPU) */
#define ASZ_TSEL_EXACT 1    /* GPU sampling + host scoring in the reference's map order   */
#define ASZ_TSEL_INJECT 2   /* caller's t_int values (block-major, round-minor)           */
/* Gradient-tuned form (no reference counterpart; it replaces the discrete
 * five-candidate scan of compress.tpp:127-142): per block, projected Newton
 * steps on the sampled residual energy sum (A t + B)^2 over [t_min, t_max],
 * the minimiser snapped to the 1/t_precision grid, and {floor, 1/4, ceil}
 * judged by the exact H1-PSNR proxy with the reference's acceptance rule.
 * The stream format is unchanged (any grid t decodes with the reference). */
#define ASZ_TSEL_GRAD 3

/* Timings (milliseconds) and diagnostics of the last asz_compress / asz_decompress. */
typedef struct asz_stats {
    double range_ms, preprocess_ms, tsel_ms, rounds_ms, huffman_ms, huffman_host_ms, total_gpu_ms, wall_ms;
    double max_err, sq_err_sum;    /* over the reconstruction, in double */
    uint64_t stream_bytes;         /* pre-Zstd Huffman stream bytes (S_h) */
    uint64_t t_count, t_fallbacks; /* #t = #blocks * layers, and fallbacks to t = 4/16 */
    uint64_t layers;
    uint64_t qb_counts[ASZ_MAX_LAYERS + 1];
    /* per-kernel device time, filled when ASZ_OPT_KERNEL_EVENTS is on */
    double pivot_ms;
    double tsel_round_ms[ASZ_MAX_LAYERS];
    double round_ms[ASZ_MAX_LAYERS];
    double hist_ms, chunk_ms, encode_ms, decode_ms;
    /* compress host timeline: from the last round launch to every codebook
     * built, the longest codebook build, then the stream layout and its upload */
    double codebook_wait_ms, codebook_max_ms, layout_ms;
    /* 1 when the call found its plan tables among the context's cached plans
     * (the last four (dims, plan, block pivot ratio, sampling count) keys) */
    uint64_t plan_cache_hit;
    /* the leading rounds that ran as block rounds (t selection and targets in
     * one launch each) and those launches' device time (ASZ_OPT_KERNEL_EVENTS;
     * also in round_ms) */
    uint64_t coarse_rounds;
    double coarse_ms;
    /* multi-field calls: the device sequence ran as a HIP graph (launched),
     * captured by this call (captured); after asz_compress_batch, the number
     * of chunks launched and captured that way */
    uint64_t graph_launched, graph_captured;
} asz_stats;

/* Options (asz_set_option).  ASZ_OPT_KERNEL_EVENTS: record a HIP event pair on
 * the context's stream around every kernel of asz_compress / asz_decompress
 * and report each kernel's device time in asz_stats. */
#define ASZ_OPT_KERNEL_EVENTS 1
/* ASZ_OPT_STREAM_ON_DEVICE: when nonzero, asz_compress leaves the pre-Zstd
 * stream in device memory and returns a device pointer in *stream (no copy to
 * the host); the caller copies it out (asz_copy_to_host) or hands it to other
 * device code.  Default 0: *stream is page-locked host memory. */
#define ASZ_OPT_STREAM_ON_DEVICE 2
/* ASZ_OPT_TIMED_ROUND: when >= 0, record the event pair around that round's
 * kernel only (asz_stats.round_ms[value]), leaving every other launch without
 * events; a negative value turns kernel events off. */
#define ASZ_OPT_TIMED_ROUND 3
/* ASZ_OPT_BATCH_FIELDS: the most fields asz_compress_batch compresses
 * together in one chunk of same-shape, same-plan fields (every per-round
 * kernel of a chunk covers all of its fields).  0 (default): automatic;
 * 1: every field on its own. */
#define ASZ_OPT_BATCH_FIELDS 4
/* ASZ_OPT_WIDE_CODES: when nonzero, the Huffman encoder takes the path it
 * keeps for codes longer than 32 bits (64-bit codebook entries ORed straight
 * into the stream) whatever the longest code is.  The stream is the same;
 * the option exists to exercise that path. */
#define ASZ_OPT_WIDE_CODES 5
/* ASZ_OPT_COARSE_ROUNDS: the most leading rounds that run as block rounds:
 * one launch per round in which a workgroup selects its block's t and then
 * predicts the block's targets (rounds with at most 16384 targets per block
 * qualify), instead of a t-selection launch and a round launch.  -1
 * (default): every leading round that qualifies; 0: none.  The stream is the
 * same either way. */
#define ASZ_OPT_COARSE_ROUNDS 6
/* ASZ_OPT_GRAPHS: when nonzero (default), a multi-field compression (the
 * batched form's chunks, GPU t selection) runs its device sequence (field
 * descriptors, pivots, every round with its t selection and folds) as a HIP
 * graph, captured the first time a (plan, field count, buffers) combination
 * is seen and replayed after that; 0: the launches are issued one by one.
 * The streams are the same either way. */
#define ASZ_OPT_GRAPHS 7
int asz_set_option(asz_ctx *ctx, int option, int64_t value);
/* Orders the context's later work after everything already enqueued on
 * `stream` (a hipStream_t of the same device; NULL: the null stream), without
 * blocking the host: callers whose inputs are produced on their own stream
 * call it before handing those device pointers over.  Every call returns with
 * the context's own work finished.  (Host integration; no reference
 * counterpart.) */
int asz_stream_wait(asz_ctx *ctx, void *stream);

/* ------------------------------------------------------------------ context */
int asz_ctx_create(int device, asz_ctx **out);
void asz_ctx_destroy(asz_ctx *ctx);
const char *asz_last_error(asz_ctx *ctx);
void asz_params_default(asz_params *p);
/* Last call's stats (valid until the next call). */
int asz_get_stats(asz_ctx *ctx, asz_stats *out);

/* Device memory helpers for callers without HIP headers (the CLI drivers).
 * No reference counterpart: the reference keeps everything in std::vector. */
int asz_device_alloc(asz_ctx *ctx, uint64_t bytes, void **d_ptr);
int asz_device_free(asz_ctx *ctx, void *d_ptr);
int asz_copy_to_device(asz_ctx *ctx, void *d_dst, const void *src, uint64_t bytes);
int asz_copy_to_host(asz_ctx *ctx, void *dst, const void *d_src, uint64_t bytes);

/* ------------------------------------------------------------------ geometry
 * generate_pivot_stride_and_block_shape_and_block_num (generate_config.tpp:51-68).
 * Host only; writes pivot_stride[3], block_shape[3], block_num[3]. */
int asz_geometry(const uint64_t dims[3], const asz_plan *plan, uint64_t block_pivot_ratio, uint64_t pivot_stride[3],
                 uint64_t block_shape[3], uint64_t block_num[3]);

/* Layout query (host only).  For round `round` of the plan, writes for every
 * target, in the round kernels' thread order, its linear field position
 * (pos) and its slot in the round's QB array (slot); with pos == NULL only
 * *count is set.  It exposes the index arithmetic the GPU kernels use for
 * block_compress_impl's loops and get_qb_index (compress.tpp:51-88,
 * asz_param.hpp:105-129) so that it can be checked without a GPU. */
int asz_round_layout(const uint64_t dims[3], const asz_plan *plan, uint64_t block_pivot_ratio, int round,
                     uint64_t *count, uint64_t *pos, uint64_t *slot);

/* ------------------------------------------------------------------ range scan
 * load_input's finite check and minmax_element (s1_compress.cpp:128-137), on the
 * GPU over d_field[n].  min/max are the first-smallest / last-largest elements
 * under the reference's FTZ/DAZ comparison, returned bit-exactly.
 * Fails with ASZ_ERR_DATA on a non-finite value or max == min. */
int asz_range_f32(asz_ctx *ctx, const float *d_field, uint64_t n, float *min_val, float *max_val);

/* Allocation-free form of asz_range_f32.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (serving loops, harnesses
 * that forbid hipMalloc).  The same scan and results; it calls no hipMalloc,
 * hipFree or other allocating routine.  Its only device memory beyond
 * d_field is the scan's result words (32 bytes), kept in a device array of
 * the library on every GPU it runs on until the process ends; their size
 * does not depend on n, so no field is too long for them.  Calls that
 * overlap on one GPU (from several contexts or threads) race on that array. */
int asz_range_f32_noalloc(asz_ctx *ctx, const float *d_field, uint64_t n, float *min_val, float *max_val);

/* ------------------------------------------------------------------ default slope
 * param_c::get_default_slope (get_default_slope.tpp:5-83): AB keys sampled on
 * the GPU, scored on the host in the reference's order.  Writes the slope and
 * the statistics the reference prints (max_invalid_ratio, fallback 0/1,
 * reason 0 none / 1 invalid_ratio / 2 invalid_difference). */
int asz_default_slope(asz_ctx *ctx, const float *d_field, const uint64_t dims[3], const asz_params *p, float min_val,
                      float max_val, float *slope, float *max_invalid_ratio, int *fallback, int *reason);

/* ------------------------------------------------------------------ compression
 * param_c::compress (compress.tpp:146-299) up to, not including, ZSTD_compress:
 * pivots, every round (per-block t, prediction, quantization, reconstruction,
 * QB serialization) and Huffman coding of the t array and the L+1 QB arrays.
 *
 * d_work holds the field (dims[0]*dims[1]*dims[2] floats, C order) and is
 * overwritten with the reconstruction (the reference's compressed_data).
 * min_val/max_val come from asz_range_f32.  tsel_mode picks the t source; with
 * ASZ_TSEL_INJECT, t_inject holds #blocks*layers t_int values.
 * On success *stream / *stream_bytes hold the pre-Zstd stream in a buffer the
 * context owns (valid until the next call on ctx): page-locked host memory,
 * or device memory when ASZ_OPT_STREAM_ON_DEVICE is set.
 * If qb_out is not NULL it receives the N QB values (pivot array then rounds);
 * if t_out is not NULL it receives the #t t_int values. */
int asz_compress(asz_ctx *ctx, float *d_work, const uint64_t dims[3], const asz_plan *plan, const asz_params *p,
                 float min_val, float max_val, int tsel_mode, const int64_t *t_inject, const uint8_t **stream,
                 uint64_t *stream_bytes, int32_t *qb_out, int64_t *t_out);

/* Range scan + compression of a resident field in one call, as load_input
 * followed by param_c::compress (s1_compress.cpp:101-140, compress.tpp:146-299):
 * the same stream as asz_range_f32 + asz_compress with the range it finds.
 * The range and the error bounds derived from it stay on the device, so the
 * host does not wait between the scan and the compression; *mn / *mx receive
 * the range.  tsel_mode: ASZ_TSEL_GPU or ASZ_TSEL_GRAD; p->default_slope is
 * used as given.  Input errors (non-finite values, max == min) are reported
 * as asz_range_f32 reports them.  (No reference counterpart as one call.) */
int asz_compress_field(asz_ctx *ctx, float *d_work, const uint64_t dims[3], const asz_plan *plan,
                       const asz_params *p, int tsel_mode, const uint8_t **stream, uint64_t *stream_bytes,
                       float *mn, float *mx);
/* Host-buffer convenience: copies field to the device, runs asz_range_f32 +
 * asz_compress, and optionally copies the reconstruction back (recon may be NULL). */
int asz_compress_host(asz_ctx *ctx, const float *field, const uint64_t dims[3], const asz_plan *plan,
                      const asz_params *p, int tsel_mode, const int64_t *t_inject, const uint8_t **stream,
                      uint64_t *stream_bytes, float *recon, int32_t *qb_out, int64_t *t_out);

/* ------------------------------------------------------------------ decompression
 * param_c::decompress (decompress.tpp:78-208) after ZSTD_decompress: parses the
 * pre-Zstd stream, Huffman-decodes every array on the GPU, and reconstructs the
 * field into d_out (dims[0]*dims[1]*dims[2] floats).  Trailing bytes after the
 * last QB array are ignored, as in the reference. */
int asz_decompress(asz_ctx *ctx, const uint8_t *stream, uint64_t stream_bytes, const uint64_t dims[3],
                   const asz_params *p, float *d_out, int32_t *qb_out, int64_t *t_out);
int asz_decompress_host(asz_ctx *ctx, const uint8_t *stream, uint64_t stream_bytes, const uint64_t dims[3],
                        const asz_params *p, float *out);
/* asz_decompress for a stream already in device memory (no reference
 * counterpart; for callers that keep streams on the GPU, e.g. the one
 * asz_compress_field leaves there under ASZ_OPT_STREAM_ON_DEVICE).  Only the
 * header bytes travel to the host (for the codebooks); the payload is read
 * where it lies.  A stream in the context's own stream buffer must start it
 * (it is then decoded in place); any other device buffer is copied first. */
int asz_decompress_device(asz_ctx *ctx, const uint8_t *d_stream, uint64_t stream_bytes, const uint64_t dims[3],
                          const asz_params *p, float *d_out);

/* ------------------------------------------------------------------ batches
 * Many fields in one call (no reference counterpart: the reference compresses
 * one field per s1_compress process, t1_start_compressors.py:482-542).  Each
 * field goes through load_input's range scan, get_default_slope and
 * param_c::compress as in asz_range_f32 + asz_default_slope + asz_compress,
 * with the same result.  Fields of one shape and plan are compressed together
 * in chunks (ASZ_OPT_BATCH_FIELDS; with ASZ_TSEL_GPU / ASZ_TSEL_GRAD), each
 * per-round kernel of a chunk covering all of its fields, and the chunks run
 * on `concurrency` worker contexts (0: 6), each with its own HIP stream and
 * host thread, so their pipelines overlap on the device.  d_fields[i]
 * (device, dims[3i..3i+2]) is overwritten with its reconstruction; plans[i]
 * is that field's plan (time steps typically share one).  On success
 * streams[i] / stream_bytes[i] hold field i's pre-Zstd stream in host memory
 * the caller releases with asz_free, or, with ASZ_OPT_STREAM_ON_DEVICE set on
 * ctx, in device memory released with asz_device_free; info (may be NULL)
 * receives per-field results.  tsel_mode: ASZ_TSEL_GPU, ASZ_TSEL_EXACT or
 * ASZ_TSEL_GRAD. */
typedef struct asz_batch_field {
    float min_val, max_val, default_slope, pad;
    double max_err, sq_err_sum;
    uint64_t t_fallbacks;
} asz_batch_field;
int asz_compress_batch(asz_ctx *ctx, uint32_t nfields, float *const *d_fields, const uint64_t *dims,
                       const asz_plan *plans, const asz_params *p, int tsel_mode, uint32_t concurrency,
                       uint8_t **streams, uint64_t *stream_bytes, asz_batch_field *info);

/* ------------------------------------------------------------------ plan generation
 * Replaces param_c::generate_config (generate_config.tpp:443-678), which the
 * reference runs when s1_compress gets no --predictor-plan
 * (s1_compress.cpp:33-36).  Reverse greedy search from the finest round: per
 * round, every candidate predictor's best t from the original field's (A, B)
 * key map, an open-loop sample compression of the round (sample stride
 * generate_config_sample_stride on two axes), the samples' Huffman + Zstd
 * size, and the largest psnr - slope * br wins.  Uses p->rel_eb,
 * eb_ratio, slope_ratio, stat_eb_ratio, h1_block_sampling_num, t_min,
 * t_max, t_precision and default_slope (from asz_default_slope).  d_field is
 * the original field in device memory (read only).  The report lists, in
 * search order (finest first), each searched round's winner (axis -1 when no
 * candidate had samples, which ends the search) and best_dist, for the
 * reference's "best_layer_predictor=..." lines, and the phase times of its
 * "Generate config breakdown" line. */
typedef struct asz_plan_report {
    uint32_t layers;                        /* rounds in the plan */
    uint32_t searched;                      /* rounds searched (layers, plus one if the search ran dry) */
    float best_dist[ASZ_MAX_LAYERS + 1];
    int8_t best_axis[ASZ_MAX_LAYERS + 1];
    int8_t best_order[ASZ_MAX_LAYERS + 1];
    double statistics_s, sample_prediction_s, lossless_estimation_s, total_s;
} asz_plan_report;
int asz_generate_plan(asz_ctx *ctx, const float *d_field, const uint64_t dims[3], const asz_params *p, float min_val,
                      float max_val, uint32_t max_layer_num, uint32_t full_predictor_search_layer_num,
                      uint32_t generate_config_sample_stride, asz_plan *plan, asz_plan_report *report);

/* ------------------------------------------------------------------ predictor family
 * Replaces param_c::select_predictor_family (predictor_family.tpp:217-251),
 * which the reference runs when s1_compress gets no --predictor-family
 * (s1_compress.cpp:44-45).  The last suffix_layers rounds of the plan at the
 * probe block pivot ratio (fine_probe_bpr when p->rel_eb <=
 * fine_eb_threshold, else coarse_probe_bpr); up to sample_blocks stratified
 * blocks (mt19937_64 seeded 0) within min(sample_point_limit,
 * floor(N * sample_fraction_limit)) points; each compressed in its own
 * context with the cubic and with the quintic interior predictor; quintic
 * wins when (quintic_psnr - cubic_psnr) - default_slope (quintic_h1 -
 * cubic_h1) > 0.  Uses p->rel_eb, pivot_eb_ratio, eb_ratio, slope_ratio,
 * stat_eb_ratio, default_slope; the probe's t sampling uses 256 samples per
 * block whatever p->h1_block_sampling_num is (the reference's probe keeps
 * that parameter at its default).  The report carries every value of the
 * reference's "Predictor family selection: ..." line (blocks == 0: the
 * point-budget fallback to cubic). */
typedef struct asz_family_opts {
    uint64_t suffix_layers;           /* predictor_family_suffix_layers (6) */
    uint64_t sample_blocks;           /* predictor_family_sample_blocks (8) */
    uint64_t sample_point_limit;      /* predictor_family_sample_point_limit (32768) */
    float sample_fraction_limit;      /* predictor_family_sample_fraction_limit (0.02) */
    float fine_eb_threshold;          /* predictor_family_fine_eb_threshold (1e-4) */
    uint64_t coarse_probe_bpr;        /* predictor_family_coarse_probe_block_pivot_ratio (3) */
    uint64_t fine_probe_bpr;          /* predictor_family_fine_probe_block_pivot_ratio (4) */
} asz_family_opts;
typedef struct asz_family_report {
    uint32_t suffix_layers, blocks;
    uint64_t sampled_block_points, sample_point_budget, probe_block_pivot_ratio, points;
    long double cubic_mse, quintic_mse, cubic_psnr, quintic_psnr, cubic_h1, quintic_h1, score_gain;
    int32_t quintic;                  /* the chosen family: 1 quintic, 0 cubic */
    double time_s;
} asz_family_report;
int asz_select_family(asz_ctx *ctx, const float *d_field, const uint64_t dims[3], const asz_plan *plan,
                      const asz_params *p, float min_val, float max_val, const asz_family_opts *opts,
                      asz_family_report *report);

/* ------------------------------------------------------------------ gradient-tuned form
 * No reference counterpart: these expose the residual-energy model behind
 * ASZ_TSEL_GRAD for tuning harnesses (the discrete scan they replace is
 * compress.tpp:90-144).
 *
 * asz_prepare_round: pivots and rounds [0, round) with the caller's t_int
 * (block-major, round-minor, #blocks * layers values; rounds >= round unused),
 * leaving d_work in the state round `round` reads (coarse points
 * reconstructed, targets original).
 *
 * asz_round_energy: per block b, E_b = sum of r_i^2 over the block's targets
 * of round `round` that use the parameterised predictor, r_i in FP32, the sum
 * in FP64 (d_E: #blocks doubles).  quintic = 0: interior 1D4R targets,
 * r = A t_b + B, d_theta = #blocks floats.  quintic = 1: 1D6R targets with
 * free weights, r = a_b (S5 - S1) + b_b (S3 - S1) + S1/2 - x0, d_theta =
 * 2 #blocks floats (a_b, b_b).
 *
 * asz_round_energy_adjoint: d_gtheta[b][j] = d_gE[b] * dE_b/dtheta_bj, FP32
 * (the vector-Jacobian product of asz_round_energy).  All pointers are device
 * pointers. */
int asz_prepare_round(asz_ctx *ctx, float *d_work, const uint64_t dims[3], const asz_plan *plan, const asz_params *p,
                      float min_val, float max_val, int round, const int64_t *t_int);
int asz_round_energy(asz_ctx *ctx, const float *d_work, const uint64_t dims[3], const asz_plan *plan,
                     const asz_params *p, int round, int quintic, const float *d_theta, double *d_E);
int asz_round_energy_adjoint(asz_ctx *ctx, const float *d_work, const uint64_t dims[3], const asz_plan *plan,
                             const asz_params *p, int round, int quintic, const float *d_theta, const float *d_gE,
                             float *d_gtheta);

/* ------------------------------------------------------------------ autotuner
 * (no reference counterpart; for the batched form)  Tile shape of one round's
 * kernel: a sweep round's segment length (targets a lane walks along the
 * round axis, at most 24) and lanes along DIM2 (1 << xsh, 6 <= xsh <= 8); a
 * DIM2 round's lattice rows per workgroup.  0 keeps the default rule.  The
 * stream and the reconstruction are the same for every tile shape.
 *
 * asz_autotune: for each round of at least 1/64 of the field's points that is
 * not a block round, times its kernel (HIP events, the median of `reps`
 * compressions of a copy of d_field, which is left untouched) under every
 * candidate tile shape and keeps the fastest; the context then uses the
 * choices for every compression of this (dims, plan, block pivot ratio,
 * sampling count).  p->default_slope is used as given.  report (optional):
 * rounds tuned, trials, and the summed kernel time of those rounds under the
 * default and the chosen shapes.
 * asz_tune_lookup: the context's choices for a key (1 found, 0 none).
 * asz_tune_save / asz_tune_load: the context's choices as a text file, one
 * line per key (load merges, replacing entries of the same key). */
typedef struct asz_round_tune {
    uint8_t seg, xsh;
    uint16_t rpw;
} asz_round_tune;
typedef struct asz_tune_report {
    uint32_t rounds_tuned, trials;
    double default_ms, tuned_ms;
} asz_tune_report;
int asz_autotune(asz_ctx *ctx, const float *d_field, const uint64_t dims[3], const asz_plan *plan,
                 const asz_params *p, uint32_t reps, asz_tune_report *report);
int asz_tune_lookup(asz_ctx *ctx, const uint64_t dims[3], const asz_plan *plan, uint64_t block_pivot_ratio,
                    uint64_t h1_block_sampling_num, asz_round_tune tune[ASZ_MAX_LAYERS]);
int asz_tune_save(asz_ctx *ctx, const char *path);
int asz_tune_load(asz_ctx *ctx, const char *path);

/* ------------------------------------------------------------------ Huffman (host side)
 * The codebook and header of one array as huffman_encoder_class<T>::preprocess_encode
 * + save build them (huffman.hpp:263-320, 650-691), from a dense histogram of
 * symbols offset..offset+range-1.  is_int64 selects the T = int64_t header
 * semantics (the t array).  Writes the header+tree bytes to hdr (capacity
 * hdr_cap), the code of each symbol to codes[range] and its length to
 * lens[range] (0 for absent symbols), and the payload bit count.  Host only. */
int asz_huffman_codebook(const uint64_t *hist, uint64_t range, int64_t offset, uint64_t count, int is_int64,
                         uint8_t *hdr, uint64_t hdr_cap, uint64_t *hdr_bytes, uint64_t *codes, uint8_t *lens,
                         uint64_t *payload_bits);

/* Compression / decompression of a whole .asz frame with host Zstd (level 3),
 * compress.tpp:294-299 / decompress.tpp:83-86.  Host only. */
int asz_zstd_compress(const uint8_t *src, uint64_t n, uint8_t **dst, uint64_t *dst_bytes);
/* The same frame at compression level `level` with `threads` host threads
 * (0: the hardware thread count, at most 16).  No reference counterpart: it
 * replaces the single ZSTD_compress call of compress.tpp:294-299 for callers
 * that want it faster.  The stream is cut into slices of whole 128 KiB
 * blocks; slice 0 is compressed with the whole stream's size pledged (so the
 * frame header is ZSTD_compress's) and every later slice on its own, and
 * their blocks are spliced into one frame when every block after slice 0 is
 * raw or RLE (stateless; Huffman output is stored raw).  Otherwise the
 * stream is compressed again in one call.  The frame carries the content
 * size decompress.tpp:83-86 reads.  *threads_used (may be NULL) receives the
 * slices used (1: one call).  Host only; release *dst with asz_free. */
int asz_zstd_compress_mt(const uint8_t *src, uint64_t n, int level, uint32_t threads, uint8_t **dst,
                         uint64_t *dst_bytes, uint32_t *threads_used);
int asz_zstd_decompress(const uint8_t *src, uint64_t n, uint8_t **dst, uint64_t *dst_bytes);
void asz_free(void *p);

#ifdef __cplusplus
}
#endif
#endif /* ASZ_GPU_H */
