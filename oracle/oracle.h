/*
 * CPU oracle for the Monte Carlo drainage-basin pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a clean-room restatement of the
 * reference's OpenMP CPU stages (jsoksane/cageo, the `drainage` program).
 * It is the checker for the HIP library and the timed CPU baseline; the
 * product never links, loads or calls it.  Only tests/, the smoke check and
 * the CPU-baseline leg of the benchmark use it.
 *
 * Every function names the reference file:line it restates.  Where the
 * reference's result depends on a third-party library that is not in the
 * reference tree (FFTW 3.3.3, Eigen JacobiSVD, thrust / cuRAND random
 * streams) the restatement follows the published mathematics and the
 * function says "parity unpinned" for the bits that library fixes.
 *
 * Build: oracle/Makefile  (g++ -O3 -fopenmp -ffp-contract=off).
 */
#ifndef CAGEO_ORACLE_H_
#define CAGEO_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes shared with the library (include/cageo_hip.h). */
#define OR_OK 0
#define OR_ERR_ARG (-1)
#define OR_ERR_RANGE (-6)

/* Label constants, FlowTracingCommon.h:18-21. */
#define OR_FLOWTRACE_OUTSIDE (-9999)
#define OR_FLOWTRACE_LAKE (-9998)

/* Philox4x32-10 block function (Salmon et al., SC'11).  Not in the
 * reference; it is the counter-based generator this build defines. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Double-precision helpers of the normal generator (exposed for tests). */
double or_rng_log(double u);
void or_rng_cossin_2pi(double t, double* c, double* s);

/* White N(mean, std^2) field, cell i drawn from Philox block i/4.
 * Follows the role of RandomFieldAlgorithmCPUOpenMP.h:47-76
 * (value = mean + r*std in float); the stream itself is this build's. */
void or_random_field(float* field, int64_t n, uint64_t seed, uint32_t realization,
                     uint32_t stream_id, float mean, float std_dev);

/* ConvolutionAlgorithm.cpp:18-172.  Writes L (Gaussian) or L*L
 * (exponential) float taps, L = 2*radius+1.  model: 0 exponential,
 * 1 Gaussian (ConvolutionCommon.h:18-21). */
int or_create_filter(int model, float practical_range, float cell_size, int radius,
                     float* out);

/* ConvolutionAlgorithmCPUOpenMP.h:70-248, in place on `field`. */
int or_convolve(float* field, int width, int height, const float* filter, int radius,
                int model);

/* RealisticDemAlgorithm.h:49-70: field = dem + field where dem != nodata,
 * then v <= 0 && v != nodata -> 1.0 (output aliases the field). */
void or_realistic_dem(const float* dem, float* field, int64_t n, float nodata_dem);

/* StreamBurningAlgorithmCPUOpenMP.h:32-82 with the offset order of
 * StreamBurningAlgorithm.h:140-167, in place on z, one-thread order. */
int or_stream_burn(float* z, const float* stream, int width, int height,
                   int bevel_radius, float nodata_dem, float nodata_stream);

/* PitFillingAlgorithmCPUOpenMP.h:68-242.  Quantizes `dem` in place and
 * writes the filled surface into `spill`. */
int or_pit_fill(float* spill, float* dem, int width, int height, float nodata_dem);

/* Second restatement of the pit fill written from the fixed-point
 * characterisation (Dijkstra order over values, a cell relays only when
 * its value survives the mm round trip).  Used to cross-check
 * or_pit_fill on small grids. */
int or_pit_fill_dijkstra(float* spill, float* dem, int width, int height, float nodata_dem);

/* FlowRoutingAlgorithmCPUOpenMP.h:53-249.  flowdir: 2 int8 per cell
 * (char2 {x, y}).  flatdist (optional, may be NULL): the internal
 * flatDist raster after flat resolution. */
void or_flow_routing(int8_t* flowdir, int32_t* flatdist, const float* filled, int width,
                     int height);

/* FlowTracingAlgorithmCPUOpenMP.h:34-134 (sequential walk + connect).
 * Returns the number of walks that hit the 100 000-step guard. */
int64_t or_flow_tracing(int32_t* labels, const int8_t* flowdir, const float* stream,
                        int width, int height);

/* CatchmentBorderAlgorithm.h:53-74 + CatchmentBorderAlgorithmCPUOpenMP.h:30-82:
 * border extraction then mc += border where mc != -1. */
void or_catchment_border(float* mc, const int32_t* labels, int width, int height);

/* One realization of main.cpp:180-283 from an injected white-noise field
 * (stages a3..a9).  `noise` is consumed (used as the work raster).  All
 * optional outputs may be NULL. */
int or_realization_from_noise(float* noise, const float* dem, const float* stream,
                              int width, int height, const float* filter, int radius,
                              int model, float nodata_dem, float nodata_stream,
                              int bevel_radius, float* mc, float* out_filled,
                              int8_t* out_flowdir, int32_t* out_labels);

/* Full realization including the random field (stage a1). */
int or_realization(const float* dem, const float* stream, int width, int height,
                   const float* filter, int radius, int model, float mean, float std_dev,
                   uint64_t seed, uint32_t realization, float nodata_dem,
                   float nodata_stream, int bevel_radius, float* mc);

#ifdef __cplusplus
}
#endif

#endif /* CAGEO_ORACLE_H_ */
