/*
 * libcageo_hip: MI355X (gfx950) implementation of the per-realization Monte
 * Carlo drainage-basin pipeline of jsoksane/cageo (`drainage`).
 *
 * Plain C ABI: device pointers and sizes, no C++ or framework types.  Each
 * stage entry point replaces one reference stage's execute()/perform*() call
 * (cited as reference-file:line, paths relative to the reference's src/).
 *
 * Conventions
 *  - Rasters are row-major W x H, index = y*W + x, resident in device memory.
 *  - Flow directions are `char2 {x, y}` (FlowDirClass.h:44): 2 int8 per cell.
 *  - Every call is enqueued on the context's HIP stream and returns after the
 *    work has completed (the reference's stages are blocking too); the return
 *    value is a CAGEO_* status code, never an exit().
 *  - A context owns the device workspace for one grid size, so no stage
 *    allocates memory per call.  One context must not be used from two host
 *    threads at once.
 */
#ifndef CAGEO_HIP_H_
#define CAGEO_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define CAGEO_OK 0
#define CAGEO_ERR_ARG (-1)           /* bad argument or shape                */
#define CAGEO_ERR_HIP (-2)           /* a HIP runtime call failed            */
#define CAGEO_ERR_NOMEM (-3)         /* device allocation failed             */
#define CAGEO_ERR_NOT_CONVERGED (-4) /* an iteration cap was reached         */
#define CAGEO_ERR_STATE (-5)         /* call order (e.g. no filter set)      */
#define CAGEO_ERR_RANGE (-6)         /* input outside the reference's domain */
#define CAGEO_ERR_ABORTED (-7)       /* a persistent launch was abandoned: the
                                        host raised the abort word, or a grid
                                        barrier waited past its time limit   */

/* ---- constants shared with the reference ------------------------------- */
#define CAGEO_FILTER_EXPONENTIAL 0 /* ConvolutionCommon.h:18-21 */
#define CAGEO_FILTER_GAUSSIAN 1
#define CAGEO_FLOWTRACE_OUTSIDE (-9999) /* FlowTracingCommon.h:18-21 */
#define CAGEO_FLOWTRACE_LAKE (-9998)
#define CAGEO_MAX_BEVEL_RADIUS 5 /* StreamBurningCommon.h:17 */

typedef struct cageo_context cageo_context;

/* Human-readable name of a status code. */
const char* cageo_status_string(int status);

/* Version string of the library build. */
const char* cageo_version(void);

/* ---- context ------------------------------------------------------------
 * Allocates the device workspace for a width x height grid on the current
 * HIP device.  `hip_stream` is a hipStream_t (NULL = the null stream).
 * Replaces the per-stage cudaMalloc/cudaFree temporaries of the reference
 * (e.g. ConvolutionAlgorithmGPUCUDA.cu:27-33, PitFillingAlgorithmGPUCUDA.cu)
 * and the ManagedData device copies (ManagedDataCUDA.h:138-213). */
int cageo_context_create(cageo_context** out, int width, int height, void* hip_stream);
int cageo_context_destroy(cageo_context* ctx);

/* ---- filter construction (host, once per run) ----------------------------
 * filterRadius = ceil((PR / cell) * 2.5) in float, main.cpp:110. */
int cageo_filter_radius(float practical_range, float cell_size);

/* ConvolutionAlgorithm<T>::createFilter, ConvolutionAlgorithm.cpp:18-45
 * (createCorrelogram :49-82, createFilterFromCorrelogram :84-148,
 * createKernelVector :150-172).  Writes L = 2*radius+1 taps for the Gaussian
 * model (separable vector) or L*L taps for the exponential model. */
int cageo_create_filter(int model, float practical_range, float cell_size, int radius,
                        float* taps_out);

/* Fast mode of the 2-D (exponential) convolution, no reference counterpart:
 * the symmetric filter's low-rank separable form f = sum_k lambda_k v_k v_k^T
 * (cyclic Jacobi eigen-decomposition in double), terms kept in decreasing
 * |lambda| until the Frobenius norm of the remainder is at most tol * |f|,
 * at most kmax.  vectors: kmax x (2r+1), values: kmax.  Returns
 * CAGEO_ERR_RANGE when kmax terms do not reach tol, CAGEO_ERR_ARG for a
 * filter whose asymmetry exceeds 1e-6 of its largest tap (below that, its
 * symmetric part is decomposed). */
int cageo_filter_low_rank(const float* taps, int radius, int kmax, double tol, float* vectors,
                          float* values, int* k_out, double* residual_out);

/* ConvolutionAlgorithm<T>::setFilterModel/setFilterRadius + the per-call
 * H2D filter copy of ConvolutionAlgorithmGPUCUDA.cu:32-33, done once. */
int cageo_set_filter(cageo_context* ctx, int model, int radius, const float* host_taps);

/* ---- stages (device pointers) ------------------------------------------ */

/* RandomFieldAlgorithm<T>::execute(field, seed), RandomFieldAlgorithm.h:82-93
 * (GPU: RandomFieldAlgorithmGPUCUDA.cu:30-88).  field = mean + std * N(0,1),
 * Philox4x32-10 keyed by `seed`, counter (cell/4, realization, stream_id). */
int cageo_random_field(cageo_context* ctx, float* d_field, uint64_t seed,
                       uint32_t realization, uint32_t stream_id, float mean, float std_dev);

/* ConvolutionAlgorithm<T>::execute(field), ConvolutionAlgorithm.h:296-312
 * (CPU order: ConvolutionAlgorithmCPUOpenMP.h:70-248).  In place, periodic
 * wrap, taps accumulated in the reference's order with every product and sum
 * rounded separately (bit-identical to the CPU implementation). */
int cageo_convolve(cageo_context* ctx, float* d_field);

/* RealisticDemAlgorithm<T,U>::execute(dem, field, field),
 * RealisticDemAlgorithm.h:49-70: field = dem + field where dem != nodata,
 * then v <= 0 && v != nodata -> 1.0.  The output aliases the field as in
 * main.cpp:200-204. */
int cageo_realistic_dem(cageo_context* ctx, const float* d_dem, float* d_field,
                        float nodata_dem);

/* StreamBurningAlgorithm<T,U>::execute(z, streams), StreamBurningAlgorithm.h:
 * 116-134 with performStreamBurning, StreamBurningAlgorithmCPUOpenMP.h:32-82.
 * In place on d_z; bevel_radius <= CAGEO_MAX_BEVEL_RADIUS. */
int cageo_stream_burn(cageo_context* ctx, float* d_z, const float* d_stream,
                      int bevel_radius, float nodata_dem, float nodata_stream);

/* PitFillingAlgorithm<T>::execute(filled, dem), PitFillingAlgorithm.h:88-98
 * (performPitFilling, PitFillingAlgorithmCPUOpenMP.h:68-242).  Quantizes
 * d_dem in place to mm as the reference does and writes the filled surface
 * to d_filled. */
int cageo_pit_fill(cageo_context* ctx, float* d_filled, float* d_dem, float nodata_dem);

/* Allocation-free variant of cageo_pit_fill.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same computation and result as cageo_pit_fill; no context.  Its workspace
 * is a file-scope device array of CAGEO_NOALLOC_SCRATCH_BYTES that every
 * device running the library holds until the process ends; width * height
 * must be at most CAGEO_NOALLOC_MAX_CELLS and the 64 x 64 tiles covering the
 * grid at most CAGEO_NOALLOC_MAX_TILES (else CAGEO_ERR_ARG).  Enqueued on
 * hip_stream (a hipStream_t; NULL = the null stream); returns when done.
 * Every call on one device shares that device's array, so calls that overlap
 * on one device race: serialise them. */
#define CAGEO_NOALLOC_MAX_CELLS (4096 * 4096)
#define CAGEO_NOALLOC_MAX_TILES 16384
#define CAGEO_NOALLOC_SCRATCH_BYTES (18 * 1024 * 1024)
int cageo_pit_fill_noalloc(float* d_filled, float* d_dem, int width, int height,
                           float nodata_dem, void* hip_stream);

/* FlowRoutingAlgorithm<T,U>::execute(flowDir, filled), FlowRoutingAlgorithm.h:
 * 69-94 (D8 FlowRoutingAlgorithmCPUOpenMP.h:53-151, flats :153-249).
 * d_flowdir: 2*W*H int8 (char2 {x, y}).  nodata_dem is accepted for the
 * interface; the reference's CPU routing does not test it. */
int cageo_flow_routing(cageo_context* ctx, int8_t* d_flowdir, const float* d_filled,
                       float nodata_dem);

/* FlowTracingAlgorithm<T,U,V>::execute(trace, flowDir, streams),
 * FlowTracingAlgorithm.h:52-69 (FlowTracingAlgorithmCPUOpenMP.h:34-134).
 * If `long_paths` is not NULL it receives the number of cells whose path to
 * its terminal is longer than the reference's 100 001-step walk guard. */
int cageo_flow_tracing(cageo_context* ctx, int32_t* d_labels, const int8_t* d_flowdir,
                       const float* d_stream, int64_t* long_paths);

/* CatchmentBorderAlgorithm<T,U>::execute(mc, trace), CatchmentBorderAlgorithm.h:
 * 53-74 (CatchmentBorderAlgorithmCPUOpenMP.h:30-82): mc += border where
 * mc != -1. */
int cageo_catchment_border(cageo_context* ctx, float* d_mc, const int32_t* d_labels);

/* ToByteArrayAlgorithm<float>::execute(bytes, values) (GPU
 * ToByteArrayAlgorithmGPUCUDA.cu:22-84, CPU ToByteArrayAlgorithmCPUOpenMP.h:
 * 32-52): bytes = (uint8) rintf(255 * ((v - min) / max)) with min and max over
 * all cells; out-of-range values wrap modulo 256 and 0/0 gives 0. */
int cageo_to_byte_array(cageo_context* ctx, uint8_t* d_bytes, const float* d_values);

/* ---- fused Monte Carlo driver (main.cpp:180-283) -------------------------
 * cageo_mc_setup binds the resident inputs and the error model; the filter
 * comes from cageo_set_filter.  cageo_run_realizations runs n realizations
 * r = first .. first+n-1 (random field seed `seed`, counter realization r)
 * and accumulates border counts into d_mc. */
int cageo_mc_setup(cageo_context* ctx, const float* d_dem, const float* d_stream,
                   float nodata_dem, float nodata_stream, int bevel_radius, float mean,
                   float std_dev);
int cageo_run_realizations(cageo_context* ctx, float* d_mc, int n, uint64_t seed,
                           uint32_t first_realization);

/* Cells, summed over the realizations of the last cageo_run_realizations,
 * cageo_realization_from_noise or cageo_flow_tracing call on this context
 * (or the last cageo_batch_run, for a tile's context), whose flow path is
 * longer than the reference's 100 001-step walk guard
 * (FlowTracingAlgorithmCPUOpenMP.h:68-75, where the reference prints
 * "Infinite loop?" and its label depends on its raster order). */
int cageo_context_long_paths(cageo_context* ctx, int64_t* out);

/* Run one realization from an injected white-noise field (stages a3..a9),
 * leaving the intermediate rasters readable for parity checks.  d_noise is
 * consumed.  Any of the optional outputs may be NULL. */
int cageo_realization_from_noise(cageo_context* ctx, float* d_noise, float* d_mc,
                                 float* d_filled_out, int8_t* d_flowdir_out,
                                 int32_t* d_labels_out);

/* The persistent kernels (pit fill, flow routing, flow tracing) wait at grid
 * barriers that need all their workgroups resident at once.  Each wait is
 * bounded: past `seconds` (default 5 s; 0 gives up at the first wait) the
 * launch is abandoned, every workgroup ends normally, and the call returns
 * CAGEO_ERR_ABORTED instead of hanging the device.  This replaces the
 * reference's host-side convergence loop over reduceAnd (Reduce.cuh:48-89),
 * which cannot deadlock but costs a host round trip per iteration.
 * Persistent launches of two contexts on concurrent streams of one device
 * can take each other's places on the device; run them one after another. */
int cageo_context_set_barrier_timeout(cageo_context* ctx, double seconds);
/* Raises (set != 0) or clears the context's abort word.  While it is raised,
 * every persistent launch of the context ends at its start, a launch already
 * running ends in its first barrier wait that stalls (about a thousand polls
 * of the barrier word), and the call returns CAGEO_ERR_ABORTED.  The only
 * call that may be made from another host thread while a call on the
 * context runs. */
int cageo_context_abort(cageo_context* ctx, int set);

/* Per-stage device time (ms, hipEvent) of the last cageo_run_realizations
 * call, summed over its realizations: random field, convolution, realistic +
 * burn, pit fill, flow routing, flow tracing, border.  `out` holds 7 floats. */
int cageo_stage_times(cageo_context* ctx, float* out, int enable);

/* ---- batched form: many tiles of uneven size x many realizations ---------
 * No reference counterpart (the reference runs one DEM per process).  Each
 * tile has its own context, error model, filter and Philox key; tile t,
 * realization r computes exactly what cageo_run_realizations does on that
 * tile alone, in every mode:
 *   CAGEO_BATCH_PER_TILE  one blocking cageo_run_realizations per tile;
 *   CAGEO_BATCH_STREAM    all tiles x realizations on one stream, one wait;
 *   CAGEO_BATCH_GRAPH     one realization of all tiles captured as a hipGraph
 *                         and replayed n times.
 * d_mc is an array of ntiles device pointers (one counter raster per tile).
 * With hip_stream NULL the batch creates its own (blocking, default-priority)
 * stream, since the NULL stream cannot be captured into a graph. */
#define CAGEO_BATCH_PER_TILE 0
#define CAGEO_BATCH_STREAM 1
#define CAGEO_BATCH_GRAPH 2
/* Convolution order (no reference counterpart).  CAGEO_CONV_EXACT (the
 * default): the reference CPU loops' order, bit-exact.  CAGEO_CONV_FAST:
 * the exponential filter's low-rank separable form with fused multiply-adds
 * (terms until the remainder is below 1e-6 of the filter, at most 32; within
 * 1e-5 of the exact convolution relative to its largest magnitude); labels
 * downstream then follow that field, not the exact one.  The Gaussian
 * filter is separable already and runs exactly in both modes. */
#define CAGEO_CONV_EXACT 0
#define CAGEO_CONV_FAST 1
int cageo_context_set_conv_mode(cageo_context* ctx, int mode);
/* Terms of the low-rank form in use (0: the exact 2-D or separable path). */
int cageo_context_conv_terms(cageo_context* ctx, int* terms);

typedef struct cageo_batch cageo_batch;
int cageo_batch_create(cageo_batch** out, int ntiles, const int* widths, const int* heights,
                       void* hip_stream);
int cageo_batch_destroy(cageo_batch* batch);
/* The per-tile context (for stage calls on one tile). */
int cageo_batch_context(cageo_batch* batch, int tile, cageo_context** ctx);
/* Inputs and error model of one tile; the filter is built from
 * (model, practical_range, cell_size) as main.cpp:110-136 does. */
int cageo_batch_setup_tile(cageo_batch* batch, int tile, const float* d_dem, const float* d_stream,
                           float nodata_dem, float nodata_stream, int bevel_radius, int model,
                           float practical_range, float cell_size, float mean, float std_dev,
                           uint64_t seed);
int cageo_batch_run(cageo_batch* batch, float* const* d_mc, int n, uint32_t first_realization,
                    int mode);
/* Autotuner: per tile, times `reps` realizations (accumulated into the
 * scratch counters d_scratch_mc) for candidate workgroup counts of the
 * tile's own persistent kernels (PER_TILE mode) and row-block heights of the
 * 2-D convolution, and keeps the fastest; then the workgroup count of the
 * all-tile persistent launches (STREAM and GRAPH modes).  out_config
 * (optional, 2 ints per tile) receives the per-tile choices.  Results do not
 * depend on these settings. */
int cageo_batch_tune(cageo_batch* batch, float* const* d_scratch_mc, int reps, int* out_config);
/* out[4]: 64x64 blocks over all tiles, and the workgroup counts of the
 * all-tile pit-fill, flow-routing and flow-tracing launches. */
int cageo_batch_info(cageo_batch* batch, int* out);
/* Raises or clears the abort word of every tile's context (see
 * cageo_context_abort); the all-tile launches read tile 0's. */
int cageo_batch_abort(cageo_batch* batch, int set);
/* Set one tile's settings by hand: persistent_workgroups (0 = full chip) and
 * conv_rows in {0 (auto), 4, 8, 16}. */
int cageo_batch_configure_tile(cageo_batch* batch, int tile, int persistent_workgroups,
                               int conv_rows);

/* Iteration counts of the last fixed-point launches on this context: pit
 * fill rounds, pit fill passes, flat-resolution passes, tracing exit-jump
 * rounds (the reference's PitFillingAlgorithm::iter-style counters).
 * `out` holds 4 ints. */
int cageo_iteration_counts(cageo_context* ctx, int32_t* out);
/* Launch settings of this context: persistent_workgroups for the persistent
 * kernels (0 = every co-resident workgroup of the device) and conv_rows, the
 * row-block height of the 2-D convolution (0 = largest that fits, 4, 8, 16).
 * Results do not depend on them. */
int cageo_context_configure(cageo_context* ctx, int persistent_workgroups, int conv_rows);

/* ---- border-probability surrogate: layer kernels ---------------------------
 * No reference counterpart (the reference has no learned part).  BF16
 * operands with FP32 accumulation on the MFMA units; activations NHWC bf16
 * (uint16 bit patterns) with channel counts that are multiples of 32.
 * All calls enqueue on `stream` (a hipStream_t) and return at once.
 *
 * 3x3 convolution, stride 1, zero padding 1, over the channel concatenation
 * [in1 | in2] (in2 may be NULL with c2 = 0); with up1 the first input is
 * read at half resolution (nearest 2x upsampling).  w: [9][cout][c1+c2]
 * bf16; bias fp32 or NULL; writes bf16 `out` (ReLU when relu) and/or fp32
 * `outf`.  With weights from cageo_nn_flip_transpose it computes the
 * gradient with respect to the input. */
int cageo_nn_conv3x3(void* stream, const uint16_t* in1, int c1, int up1, const uint16_t* in2,
                     int c2, const uint16_t* w, const float* bias, uint16_t* out, float* outf,
                     int n, int h, int w_, int cout, int relu);
/* The forward convolution (bf16 output, no fp32 output) with consumers of
 * its output fused into the launch, each giving the same bits as its own
 * kernel run after the convolution:
 *   pool (or null): cageo_nn_avgpool2 of the output, [n][h/2][w/2][cout];
 *   head_w, head_b, prob (or null): cageo_nn_head of the output into prob
 *     ([n][h][w] fp32, the sigmoid when head_sigmoid); out and pool must
 *     then be null (the output never reaches memory).
 * out may be null when the output itself is not needed.  Only where
 * cageo_nn_conv3x3_fusable(h, w_, cout, head) returns 1 (w_ % 64 == 0,
 * h % 4 == 0, cout % 32 == 0; with the head, cout 32 or 64: one workgroup
 * holds all of a pixel's channels); CAGEO_ERR_ARG otherwise.  No reference
 * counterpart (surrogate inference). */
int cageo_nn_conv3x3_fused(void* stream, const uint16_t* in1, int c1, int up1, const uint16_t* in2,
                           int c2, const uint16_t* w, const float* bias, uint16_t* out,
                           uint16_t* pool, const float* head_w, const float* head_b, float* prob,
                           int head_sigmoid, int n, int h, int w_, int cout, int relu);
int cageo_nn_conv3x3_fusable(int h, int w_, int cout, int head);
/* dw[tap][co][ci] += sum_p g[p][co] * in[p + offset(tap)][ci] (fp32 dw). */
int cageo_nn_conv3x3_wgrad(void* stream, const uint16_t* in1, int c1, int up1,
                           const uint16_t* in2, int c2, const uint16_t* g, float* dw, int n, int h,
                           int w_, int cout);
/* wt[8 - tap][ci][co] = w[tap][co][ci]. */
int cageo_nn_flip_transpose(void* stream, const uint16_t* w, uint16_t* wt, int cout, int cin);
int cageo_nn_avgpool2(void* stream, const uint16_t* in, uint16_t* out, int n, int h, int w_, int c);
/* dx (fp32, h x w) += gradient of the 2x2 average pool from dy (h/2 x w/2). */
int cageo_nn_avgpool2_back(void* stream, const float* dy, float* dx, int n, int h, int w_, int c);
/* dx (half resolution) += 2x2 sums of channels [coff, coff+c) of dy. */
int cageo_nn_upsample2_back(void* stream, const float* dy, float* dx, int n, int h, int w_, int c,
                            int ctot, int coff);
/* dx[p][0..c) (+)= dy[p][coff..coff+c) */
int cageo_nn_slice_add(void* stream, const float* dy, float* dx, long long p, int c, int ctot,
                       int coff, int accumulate);
/* g = bf16(dy * (act > 0)) */
int cageo_nn_relu_back(void* stream, const float* dy, const uint16_t* act, uint16_t* g,
                       long long n);
/* db[c] += sum_p g[p][c] */
int cageo_nn_bias_grad(void* stream, const uint16_t* g, float* db, long long p, int c);
/* Both in one pass: g = bf16(dy * (act > 0)) over p pixels of c channels
 * (c divides 256), db[c] += sum_p g[p][c]. */
int cageo_nn_relu_back_bias(void* stream, const float* dy, const uint16_t* act, uint16_t* g,
                            long long p, int c, float* db);
/* logit[p] = b + sum_c w[c] act[p][c]; with sigmoid != 0 the probability
 * 1 / (1 + exp(-logit)) is written instead. */
int cageo_nn_head(void* stream, const uint16_t* act, const float* w, const float* b, float* logit,
                  long long p, int c, int sigmoid);
/* q = sigmoid(logit); loss += mean BCE(q, target); d(act), dw, db of the head. */
int cageo_nn_head_back(void* stream, const uint16_t* act, const float* w, const float* logit,
                       const float* target, float* q, float* dact, float* dw, float* db,
                       float* loss, long long p, int c);
/* Adam on fp32 master weights p (bias-corrected, step >= 1); refreshes the
 * bf16 copy pb when not NULL. */
int cageo_nn_adam(void* stream, float* p, const float* g, float* m, float* v, uint16_t* pb,
                  long long n, float lr, float b1, float b2, float eps, int step);
int cageo_nn_to_bf16(void* stream, const float* x, uint16_t* y, long long n);

/* ---- synthetic inputs (host) -----------------------------------------------
 * Seeded DEM: 100 m base + 0.5 % tilt along +x + 40 Gaussian hills
 * (height U(5,60) m, sigma U(50,400) m) + N(0, 0.3^2) noise; streams where
 * D8 flow accumulation on the filled 5x5-smoothed DEM exceeds acc_threshold
 * cells, float ids >= 1 per stream link, nodata_stream elsewhere. */
int cageo_synth_dem(int width, int height, uint64_t seed, float cell_size,
                    int acc_threshold, float nodata_stream, float* dem_out,
                    float* stream_out);

/* ---- border-probability surrogate: inference and check ----------------
 * No reference counterpart (the reference only runs the Monte Carlo).  A
 * U-Net on the cageo_nn_* kernels maps the z-scored DEM, the stream mask,
 * RMSE / 3 m and practical range / 200 m to the per-cell probability that
 * a realization puts a catchment border there (the Monte Carlo's count /
 * realizations).  Weights: the `.bin` files cageo_amd/surrogate.py writes
 * (magic "CGSURR01"; convolution weights bf16, biases float32).  Calls on
 * one surrogate must not overlap (it owns one workspace). */
typedef struct cageo_surrogate cageo_surrogate;
int cageo_surrogate_create(cageo_surrogate** out, const char* weights_path, void* hip_stream);
int cageo_surrogate_destroy(cageo_surrogate* s);
/* Forward pass of prepared inputs: bf16 NHWC [h][w][32] (h, w multiples of
 * 8) -> probabilities [h][w] (float). */
int cageo_surrogate_forward(cageo_surrogate* s, const uint16_t* d_x, int h, int w, float* d_prob);
/* DEM + stream raster (W x H, any size) + error model -> probabilities
 * [H][W]: input planes built on the device (mean and population standard
 * deviation of the valid DEM cells), padded to multiples of 8, cropped. */
int cageo_surrogate_predict(cageo_surrogate* s, const float* d_dem, const float* d_stream, int W,
                            int H, float nodata_dem, float nodata_stream, float rmse,
                            float practical_range, float* d_prob);
/* Monte Carlo check of a probability map d_prob (W x H floats of the
 * context's grid), no reference counterpart.  k >= 2 realizations of `ctx`
 * (set up with cageo_mc_setup and cageo_set_filter; counts into d_scratch,
 * W x H floats) give f = count / k per cell, and per window
 *   bias = mean(f - q)                             (expectation mean(p - q))
 *   rms  = sqrt(max(mean((f - q)^2 - f (1 - f) / (k - 1)), 0))
 *          (the mean of an unbiased estimate of (p - q)^2 per cell)
 * with p the true border probability: neither is inflated by the sampling
 * noise of a small k.  accepted = |bias| <= bias_tol and rms <= rms_tol.
 * Windows are `window` x `window` cells in row-major order, ragged at the
 * right and bottom edges; window <= 0 makes the whole raster one window.
 * accepted, bias, rms (each optional) hold one entry per window:
 * cageo_check_windows gives their count. */
int cageo_check_windows(cageo_context* ctx, int window, int* nwx, int* nwy);
int cageo_check_probability_map(cageo_context* ctx, const float* d_prob, int k, uint64_t seed,
                                int window, double bias_tol, double rms_tol, float* d_scratch,
                                int* accepted, double* bias, double* rms);
/* Refinement of the windows that failed the check: n realizations of the
 * Monte Carlo over the whole DEM (a window's labels depend on flow paths that
 * leave it) into d_scratch, whose frequency count / n replaces d_prob in every
 * window with accepted[w] == 0 (host array, one flag per window as above).
 * Nothing runs when every window was accepted. */
int cageo_refine_probability_map(cageo_context* ctx, float* d_prob, int window,
                                 const int* accepted, int n, uint64_t seed, float* d_scratch);
/* The check above on the whole raster (the surrogate's own map). */
int cageo_surrogate_check(cageo_surrogate* s, cageo_context* ctx, const float* d_prob, int k,
                          uint64_t seed, double bias_tol, double rms_tol, float* d_scratch,
                          int* accepted, double* bias, double* rms);

/* ---- border-probability surrogate: training --------------------------------
 * No reference counterpart (the reference has no learned part).  A trainer
 * holds the network's fp32 master weights, their bf16 copies, gradients
 * and Adam moments (beta 0.9 / 0.999, eps 1e-8) for base `base` (a multiple
 * of 32), and the workspace of one batch shape.  Initial weights: the
 * weights file `init_weights` (surrogate format; NULL: He-normal from
 * `seed`, zero biases).  Inputs are fp32 NHWC [tiles][h][w][4] (z-scored
 * DEM, stream mask 0/1, RMSE / 3 m, practical range / 200 m; h, w multiples
 * of 8), targets the Monte Carlo frequencies [tiles][h][w].  The loss is
 * the mean binary cross-entropy of the sigmoid output.  Calls on one
 * trainer must not overlap. */
typedef struct cageo_trainer cageo_trainer;
int cageo_trainer_create(cageo_trainer** out, int base, uint64_t seed, const char* init_weights,
                         void* hip_stream);
int cageo_trainer_destroy(cageo_trainer* t);
/* Loss and parameter gradients (replaced, not added) of the batch of the n
 * tiles of d_x / d_target, without an update. */
int cageo_trainer_gradients(cageo_trainer* t, const float* d_x, const float* d_target, int n,
                            int h, int w, float* loss);
/* The same, then one Adam step with learning rate lr. */
int cageo_trainer_step(cageo_trainer* t, const float* d_x, const float* d_target, int n, int h,
                       int w, float lr, float* loss);
/* Steps first_step .. first_step + nsteps - 1 of a run of total_steps Adam
 * steps over the ntiles tiles of d_x / d_y: each step's batch is `batch`
 * tiles drawn with replacement from (seed, step number), each under a
 * random dihedral transform (quarter turns only for square tiles); the
 * learning rate rises linearly to lr over the first 2 % of the steps (at
 * most 200), then follows a half cosine from lr over total_steps, not
 * below lr * lr_floor.  A run split into pieces computes what one call
 * would.  loss_hist (host, nsteps floats, may be NULL) gets each step's
 * loss. */
int cageo_surrogate_train(cageo_trainer* t, const float* d_x, const float* d_y, int ntiles, int h,
                          int w, int batch, float lr, float lr_floor, int first_step, int nsteps,
                          int total_steps, uint64_t seed, float* loss_hist);
/* Host copy of parameter `name` (what 0) or of its gradient (what 1);
 * n must be its size. */
int cageo_trainer_read(cageo_trainer* t, const char* name, int what, float* host, long long n);
/* Weights file (the format cageo_surrogate_create reads). */
int cageo_trainer_save(cageo_trainer* t, const char* path);

#ifdef __cplusplus
}
#endif

#endif /* CAGEO_HIP_H_ */
