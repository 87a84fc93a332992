/* C ABI of libcp3hip.so -- the MI355X (gfx950) propagation library.
 *
 * It replaces the device side of the reference's single-GPU propagator
 * (cudapropagator.cuh:39-251, class CudaPropagatorSingle) and the CUDA kernel it
 * launches (physics.hpp:583-742).  Host-side set-up -- the PMNS matrix, the mass
 * differences, the density flip, coslimit and maxlayers -- stays with the caller's
 * Propagator<FLOAT_T> (include/cudaprob3/propagator.hpp, following propagator.hpp:117-307),
 * which hands the results across this boundary as plain arrays.
 *
 * Conventions: every function returns CP3H_OK (0) or a nonzero status; the message of
 * the last failure on the calling thread is cp3h_last_error().  Arrays are host
 * pointers unless stated otherwise.  A handle is not thread-safe; distinct handles are
 * independent (all parameters travel per launch, no process-global device state).
 * Results are laid out SoA: result[t * n_cosines * n_energies + ic * n_energies + ie],
 * t = in*3 + out (types.hpp:23-27), as the reference GPU path writes them
 * (physics.hpp:689-691).
 */
#ifndef CP3HIP_H
#define CP3HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cp3h_propagator cp3h_propagator;

enum cp3h_status {
    CP3H_OK = 0,
    CP3H_ERR_ARGUMENT = 1, /* bad size, index, pointer or enum */
    CP3H_ERR_STATE = 2,    /* call order violated (e.g. compute before inputs are set) */
    CP3H_ERR_DEVICE = 3,   /* no GPU / bad device id (cudapropagator.cuh:56, 64) */
    CP3H_ERR_HIP = 4       /* a HIP runtime call failed; message holds hipGetErrorString */
};

/* Arithmetic of the per-pair kernels. */
enum cp3h_precision {
    /* FP64 throughout, the reference's per-pair Barger step with its rounding order
       (results within 1e-12 of the reference CpuPropagator<double>). */
    CP3H_FP64_PLAIN = 0,
    /* FP64, eigen-decomposition hoisted to one per (shell, energy) (spectrum table). */
    CP3H_FP64 = 1,
    /* FP32 bulk arithmetic, FP64 eigenvalues and FP64 phase reduction; float results. */
    CP3H_MIXED = 2
};

/* Number of visible GPUs.  cudapropagator.cuh:52-54. */
int cp3h_device_count(int* count);

/* cudapropagator.cuh:50-80 (CudaPropagatorSingle(int id, int n_cosines, int n_energies)).
   Allocates the device lists and the result buffer on `device`, creates the stream. */
int cp3h_create(int device, int n_cosines, int n_energies, int precision, cp3h_propagator** out);

/* cudapropagator.cuh:91-94. Frees everything; NULL is accepted. */
int cp3h_destroy(cp3h_propagator* h);

/* cudapropagator.cuh:148-153 (setEnergyList).  n must equal n_energies; GeV. */
int cp3h_set_energies_f64(cp3h_propagator* h, const double* energies, int n);
int cp3h_set_energies_f32(cp3h_propagator* h, const float* energies, int n);

/* cudapropagator.cuh:155-159 and 184-188 (setCosineList + setMaxlayers).  maxlayers[i] is
   the number of shells crossed at cosines[i] (propagator.hpp:301-307). */
int cp3h_set_cosines_f64(cp3h_propagator* h, const double* cosines, const int* maxlayers, int n);
int cp3h_set_cosines_f32(cp3h_propagator* h, const float* cosines, const int* maxlayers, int n);

/* Update maxlayers only (after a density change re-derives them, propagator.hpp:160). */
int cp3h_set_maxlayers(cp3h_propagator* h, const int* maxlayers, int n);

/* cudapropagator.cuh:132-146 (setDensity): radii in km in DESCENDING order (radii[0] =
   Earth radius, last = 0) and the density (g/cm^3) of the shell whose outer radius is
   radii[i], as the base class leaves them after its flip (propagator.hpp:141-147). */
int cp3h_set_density_f64(cp3h_propagator* h, const double* radii_desc, const double* rhos_desc, int n_rows);
int cp3h_set_density_f32(cp3h_propagator* h, const float* radii_desc, const float* rhos_desc, int n_rows);

/* physics.hpp:130-194 (setMixMatrix + setMassDifferences): U row-major with (re, im)
   interleaved (18 values), DM row-major 3x3.  The vacuum root order of
   physics.hpp:198-253 is derived here on the host in the same precision. */
int cp3h_set_params_f64(cp3h_propagator* h, const double* U18, const double* DM9);
int cp3h_set_params_f32(cp3h_propagator* h, const float* U18, const float* DM9);

/* cudapropagator.cuh:191-220 (calculateProbabilitiesAsync): enqueue the propagation on
   the handle's stream and return.  antineutrino: 0 or 1 (types.hpp:29).
   production_height_cm: the base class's ProductionHeightinCentimeter. */
int cp3h_compute_async(cp3h_propagator* h, int antineutrino, double production_height_cm);

/* cudapropagator.cuh:223-226 (waitForCompletion). */
int cp3h_wait(cp3h_propagator* h);

/* cudapropagator.cuh:162-165 (calculateProbabilities) = compute_async + wait. */
int cp3h_compute(cp3h_propagator* h, int antineutrino, double production_height_cm);

/* cudapropagator.cuh:229-235 (getResultFromDevice) with the pinned result buffer of
   cudapropagator.cuh:73: copies all 9*n_c*n_e results (SoA, the handle's element type: 8 bytes
   for the FP64 modes, 4 for CP3H_MIXED) into a pinned host buffer the handle owns (allocated on
   the first call) and returns it.  Valid until the next fetch or the handle's destruction. */
int cp3h_fetch_results(cp3h_propagator* h, const void** host_ptr, int* elem_bytes);

/* The same into a caller array, converting to the requested type (through the pinned buffer). */
int cp3h_copy_results_f64(cp3h_propagator* h, double* dst, size_t count);
int cp3h_copy_results_f32(cp3h_propagator* h, float* dst, size_t count);

/* Device-resident results for downstream kernels: device pointer, element size in bytes
   (8 for the FP64 modes, 4 for CP3H_MIXED).  Valid until the handle is destroyed. */
int cp3h_device_results(cp3h_propagator* h, const void** ptr, int* elem_bytes);

/* The handle's hipStream_t (as void*), for callers that order their own work after it. */
int cp3h_stream(cp3h_propagator* h, void** stream);

/* Run the handle's work on the caller's hipStream_t (as void*, created on the handle's
   device; NULL returns to the handle's own stream).  The caller keeps ownership. */
int cp3h_set_stream(cp3h_propagator* h, void* stream);

/* Copy the results (SoA, the handle's element type) to device memory of the same size. */
int cp3h_copy_results_device(cp3h_propagator* h, void* dst_device, size_t count);

/* sum over all 9*n_c*n_e results of (result - data)^2, accumulated in FP64 on the device;
   data is a device float array in the results' SoA layout (CP3H_MIXED handles).  Orders
   after the last compute on the handle's stream and waits for the scalar. */
int cp3h_sqdiff_results_f32(cp3h_propagator* h, const float* data_device, double* result);

/* calculateProbabilities fused with that comparison (no counterpart in the reference): the FP64
   sum over all 9*n_c*n_e probabilities of (P - data)^2 for the current parameters, computed
   by the mixed kernel without storing P (the handle's results are left as they were);
   CP3H_MIXED handles; waits for the result.  The pair-kernel time is reported by
   cp3h_last_compute_ms. */
int cp3h_compute_sqdiff_f32(cp3h_propagator* h, int antineutrino, double production_height_cm,
                            const float* data_device, double* result);

/* GPU time of the last compute, from HIP events on the handle's stream (waits for it):
   total_ms covers all of its kernels, pairs_ms only the per-pair propagation kernel
   (for CP3H_FP64_PLAIN the two are the same kernel).  Either pointer may be NULL. */
int cp3h_last_compute_ms(cp3h_propagator* h, float* total_ms, float* pairs_ms);
/* Timing of the computations (no counterpart in the reference): with enable != 0 every compute
   records HIP events around its kernels for cp3h_last_compute_ms; off by default (each event
   record adds a gap of several microseconds between the kernels). */
int cp3h_set_timing(cp3h_propagator* h, int enable);

/* Benchmark support (no counterpart in the reference; the reference times only wall clock,
   example/main.cpp:130-139): launches the computation's set-up kernels `reps` times back to back,
   then its per-pair kernel `reps` times back to back, between HIP events on the handle's stream,
   and returns the average duration of one launch of each (setup_ms is 0 for CP3H_FP64_PLAIN,
   which has no set-up).  Back-to-back launches hide the gaps that an event pair around a single
   launch would include.  data_device != NULL times the likelihood variant of
   cp3h_compute_sqdiff_f32 (CP3H_MIXED) instead of the storing kernel.  Waits for completion; the
   results are those of one computation.  Either pointer may be NULL. */
int cp3h_time_kernels(cp3h_propagator* h, int antineutrino, double production_height_cm, const float* data_device,
                      int reps, float* setup_ms, float* pairs_ms);

/* Roofline calibration (no counterpart in the reference): the sustained rate, in TFLOP/s, of
   plain (non-packed) FP32 and FP64 vector fused multiply-adds on `device`: independent FMA
   chains in every lane, eight wavefronts per SIMD; the best of five launches of about 2 ms, each
   between HIP events, after 40 untimed launches that bring the clocks up (about 0.15 s in all).  It is the ceiling of kernels whose arithmetic issues
   one FMA per instruction, as this library's do (packed FP32 is disabled), beside the data
   sheet's FP32 figure, which counts packed FP32.  Allocates and frees a small device buffer;
   either pointer may be NULL to skip that precision. */
int cp3h_measure_valu_peak(int device, float* fp32_tflops, float* fp64_tflops);

/* Message of the last failing call on this thread ("" if none). */
const char* cp3h_last_error(void);

/* Allocation-free propagation.  No counterpart in the reference: for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc).  It is cp3h_compute_async of a CP3H_MIXED handle without the handle: the
   grid, the Earth model and the results are caller device memory, the parameters are host
   arrays as for cp3h_set_params_f64, and the kernels' scratch (layer lengths and the spectrum
   table) lives in one static device array of 128 MiB per GPU owned by the library.
     cosines_device[n_cosines], maxlayers_device[n_cosines] (each in [0, n_rows]; entries outside
     are clamped into it, so no read leaves the arrays, but such rows' results are meaningless),
     energies_device[n_energies] (GeV), radii_desc_device / rhos_desc_device[n_rows] (descending,
     as for cp3h_set_density_f64), out_device: float SoA [9][n_cosines][n_energies].
   Enqueues two kernels on `stream` (a hipStream_t as void*, NULL = the null stream) and returns;
   calls no hipMalloc, hipFree or other allocating routine.  Limit: 8 n_cosines (n_rows + 1) +
   96 (n_rows + 1) n_energies_padded bytes <= 128 MiB (n_energies_padded = n_energies rounded up
   to 64; each of the three parts rounded up to 256 bytes), e.g. 8192^2 on 100 shells; larger grids are refused with CP3H_ERR_ARGUMENT.  All calls on one GPU share its
   scratch: calls whose device work overlaps (different streams, no ordering) race; calls on one
   stream are ordered.  The first call on a device looks the scratch's address up (make it
   outside graph capture). */
int cp3h_propagate_mixed_noalloc(int device, const double* cosines_device, const int* maxlayers_device,
                                 int n_cosines, const double* energies_device, int n_energies,
                                 const double* radii_desc_device, const double* rhos_desc_device, int n_rows,
                                 const double* U18, const double* DM9, int antineutrino,
                                 double production_height_cm, float* out_device, void* stream);

/* ---- Batched parameter scans ------------------------------------------------------------
 * No counterpart in the reference, whose propagators hold one hypothesis, one grid and one
 * Earth model each: here many parameter sets, each with its own grid and Earth model, are
 * propagated in one persistent launch (mixed precision, float results), with the set-up
 * kernel and the pair kernel optionally replayed as a captured HIP graph.  Each set's inputs
 * follow the single-propagator conventions above (host set-up by Propagator<FLOAT_T>). */
typedef struct cp3h_batch cp3h_batch;

/* n_sets sets; set i has n_cosines[i] x n_energies[i] pairs and an Earth model of n_rows[i]
   rows.  All device memory is allocated here. */
int cp3h_batch_create(int device, int n_sets, const int* n_cosines, const int* n_energies, const int* n_rows,
                      cp3h_batch** out);
int cp3h_batch_destroy(cp3h_batch* b);

/* Grid of set i: cosines with their maxlayers, energies (GeV). */
int cp3h_batch_set_grid_f64(cp3h_batch* b, int set, const double* cosines, const int* maxlayers,
                            const double* energies);
/* Earth model of set i, descending radii as for cp3h_set_density_f64. */
int cp3h_batch_set_density_f64(cp3h_batch* b, int set, const double* radii_desc, const double* rhos_desc);
/* Hypothesis of set i: U (18), DM (9), neutrino type, production height in cm.  May be
   changed between computations (a parameter-scan step). */
int cp3h_batch_set_params_f64(cp3h_batch* b, int set, const double* U18, const double* DM9, int antineutrino,
                              double production_height_cm);
/* The same for sets first .. first+count-1 in one call: U (18 per set), DM (9 per set),
   neutrino types and heights (one per set). */
int cp3h_batch_set_params_many_f64(cp3h_batch* b, int first, int count, const double* U18s, const double* DM9s,
                                   const int* antineutrino, const double* production_height_cm);

/* Launch configuration of the persistent pair kernel: waves per workgroup (4, 8, 16) and
   workgroups per CU (1-16).  cp3h_batch_autotune times every combination on the current
   inputs and keeps the fastest; returns them and the best time of the pair kernel. */
int cp3h_batch_config(cp3h_batch* b, int waves_per_group, int groups_per_cu);
int cp3h_batch_autotune(cp3h_batch* b, int* waves_per_group, int* groups_per_cu, float* best_ms);

/* 1: computations replay a HIP graph (parameter upload, set-up kernel, pair kernel)
   captured on first use and re-captured after a configuration change; 0: plain launches. */
int cp3h_batch_use_graph(cp3h_batch* b, int enable);

int cp3h_batch_compute_async(cp3h_batch* b);
int cp3h_batch_wait(cp3h_batch* b);
int cp3h_batch_compute(cp3h_batch* b);

/* GPU time of the last computation (waits for it): all kernels, and the pair kernel alone
   (pairs_ms is -1 when the computation was a graph replay). */
int cp3h_batch_last_ms(cp3h_batch* b, float* total_ms, float* pairs_ms);
/* Benchmark support: the set-up kernel `reps` times back to back, then the persistent pair
   kernel `reps` times back to back (each after a one-workgroup reset of its work queues, as the
   set-up does), between HIP events; returns the average duration of one launch of each.  Uses
   the pending hypotheses; waits for completion. */
int cp3h_batch_time_kernels(cp3h_batch* b, int reps, float* setup_ms, float* pairs_ms);

/* Results of set i, SoA [t][ic][ie] float, count = 9 * n_cosines[i] * n_energies[i] (copied
   densely).  On the device each cosine row starts a 128-byte line: P(t, ic, ie) is at
   ptr[t * plane_stride + ic * row_stride + ie], row_stride = n_energies[i] rounded up to 32, so
   that the kernel's row stores never share a line with another row (no counterpart in the
   reference, whose batches do not exist). */
int cp3h_batch_copy_results_f32(cp3h_batch* b, int set, float* dst, size_t count);
int cp3h_batch_device_results(cp3h_batch* b, int set, const float** ptr);
int cp3h_batch_result_strides(cp3h_batch* b, int set, long long* plane_stride, int* row_stride);

/* Energy-bin averages of sets first .. first+count-1 whose energy lists are `nodes`
   quadrature nodes per bin (bin-major): out_device receives, set after set, the SoA array
   [9][n_cosines][n_energies / nodes] of sum_g weights[g] * P(node g).  Orders after the last
   computation on the batch's stream and waits. */
int cp3h_batch_bin_average_f32(cp3h_batch* b, int first, int count, int nodes, const float* weights,
                               float* out_device);

/* Binned mode (no counterpart in the reference): with nodes > 0 every set's energy list must
   be `nodes` quadrature nodes per bin (bin-major, nodes a power of two <= 32) and each
   computation stores only the bin averages sum_g weights[g] * P(node g), for set i at
   out_device[offsets[i] + t * plane_strides[i] + ic * row_strides[i] + bin] (t = in*3+out), so
   many sets can fill disjoint blocks of one larger [9][rows][bins] array.  No unbinned results
   are kept.  nodes = 0 returns to unbinned results. */
int cp3h_batch_set_binning(cp3h_batch* b, int nodes, const float* weights, float* out_device,
                           const long long* offsets, const long long* plane_strides, const int* row_strides);

/* Per-row energy limits (no counterpart in the reference): cosine row ic of `set` computes and
   stores only its first limits[ic] energies (in binned mode: the bins starting below it);
   limits = NULL removes them.  Lets one set (one spectrum table) cover rows that need different
   energy ranges. */
int cp3h_batch_set_row_limits(cp3h_batch* b, int set, const int* limits);

/* ---- Surrogate network --------------------------------------------------------------------
 * No counterpart in the reference.  A 9 -> 256 x 5 (GELU, tanh form) -> 9 MLP with BF16
 * weights and FP32 accumulation on the matrix cores, approximating energy-bin-averaged
 * probabilities from (cosZ, log10 E, theta12, theta13, theta23, deltaCP, dm21^2, dm32^2,
 * nubar flag).  Activations enter each product as two bf16 terms (hi + lo, about 16
 * significant bits).  Inputs are device float arrays [n][9]; outputs device float [n][9] in
 * the order t = in*3 + out.  Parameters (fp32, cp3h_mlp_param_count of them) are layer by
 * layer: W_l row-major [outputs][inputs], then b_l. */
typedef struct cp3h_mlp cp3h_mlp;

int cp3h_mlp_param_count(size_t* n);
/* max_batch: rows per internal launch (multiple of 256); training: also keep the backward
   buffers, gradients and Adam state. */
int cp3h_mlp_create(int device, int max_batch, int training, cp3h_mlp** out);
int cp3h_mlp_destroy(cp3h_mlp* h);
int cp3h_mlp_set_stream(cp3h_mlp* h, void* stream);
/* inputs are mapped to (x - shift) * scale before the first layer */
int cp3h_mlp_set_scaling(cp3h_mlp* h, const float* shift9, const float* scale9);
/* sets the fp32 parameters (and resets the optimizer state) / reads them back */
int cp3h_mlp_set_params(cp3h_mlp* h, const float* params, size_t n);
int cp3h_mlp_get_params(cp3h_mlp* h, float* params, size_t n);
/* inference of n rows (asynchronous on the network's stream) */
int cp3h_mlp_forward(cp3h_mlp* h, const float* x_device, long long n, float* y_device);
/* inference kernel: CP3H_MLP_FUSED4 / CP3H_MLP_FUSED8 run all layers in one launch per 128
   rows (4 or 8 wavefronts per workgroup), CP3H_MLP_FUSED64 per 64 rows (two workgroups per CU);
   FUSED4 is the default; CP3H_MLP_LAYERED runs the training path's per-layer GEMMs in chunks of
   max_batch rows (row mode only; grid mode then uses FUSED4). */
enum { CP3H_MLP_LAYERED = 0, CP3H_MLP_FUSED4 = 1, CP3H_MLP_FUSED8 = 2, CP3H_MLP_FUSED64 = 3 };
int cp3h_mlp_set_forward_kernel(cp3h_mlp* h, int kernel);
/* Inference on n_hyp hypotheses over one cosZ x energy-bin grid (fused kernel): the inputs
   of point (h, ic, ib) are cos[ic], log10e[ib] (log10 of the bin centre / GeV) and hyp[h][0..6]
   (theta12, theta13, theta23, deltaCP, dm21^2, dm32^2, nubar flag); outputs go to
   y[h][t][ic][ib], the propagators' SoA layout.  skip_below (device [n_hyp][n_cos], may be
   NULL): points with ib < skip_below[h][ic] are left unwritten (bins another kernel fills,
   possibly at the same time). */
int cp3h_mlp_forward_grid(cp3h_mlp* h, const float* cos_device, int n_cos, const float* log10e_device, int n_bins,
                          const float* hyp_device, int n_hyp, const int* skip_below_device, float* y_device);
/* The same over listed tiles only: tiles_device[0 .. n_tiles) are indices t of 128-row tiles of
   the h-major row order (rows 128 t .. 128 t + 127, each < n_hyp * n_cos * n_bins, no repeats);
   points of other tiles, and points below skip_below (may be NULL), are left unwritten.  Uses
   FUSED8 when that kernel is selected, else FUSED4. */
int cp3h_mlp_forward_grid_tiles(cp3h_mlp* h, const float* cos_device, int n_cos, const float* log10e_device,
                                int n_bins, const float* hyp_device, int n_hyp, const int* skip_below_device,
                                const long long* tiles_device, long long n_tiles, float* y_device);
/* gradients of the mean squared error over the 9 outputs of one batch (M a multiple of 256,
   <= max_batch), copied to host memory in the parameter layout, and the loss */
int cp3h_mlp_gradients(cp3h_mlp* h, const float* x_device, const float* t_device, int M, float* grads, size_t n,
                       float* loss);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on one batch; loss may be NULL (no wait) */
int cp3h_mlp_train_step(cp3h_mlp* h, const float* x_device, const float* t_device, int M, float lr, float* loss);
int cp3h_mlp_sync(cp3h_mlp* h);

#ifdef __cplusplus
}
#endif

#endif /* CP3HIP_H */
