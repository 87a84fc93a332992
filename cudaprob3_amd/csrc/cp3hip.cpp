// Host side of libcp3hip.so: the C ABI declared in include/cp3hip.h.
//
// One handle = one single-GPU propagator (the device half of the reference's
// CudaPropagatorSingle, cudapropagator.cuh:39-251): device copies of the grid and the
// Earth model, the result buffer, per-(shell, energy) tables, one HIP stream and two
// timing events.  Oscillation parameters live in the handle on the host and reach the
// kernels by value at launch, so there is no process-global device state.
#include "../../include/cp3hip.h"
#include "common.hpp"
#include "host_util.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

using cp3host::fail;
using cp3host::vacuum_order;

struct cp3h_propagator {
    int device = 0, n_c = 0, n_e = 0, precision = CP3H_FP64_PLAIN, ne_pad = 0, rows = 0;
    hipStream_t stream = nullptr;      // the stream work is enqueued on
    hipStream_t own_stream = nullptr;  // created and destroyed by the handle
    hipEvent_t ev_start = nullptr, ev_pairs = nullptr, ev_stop = nullptr;
    void* h_results = nullptr;         // pinned host copy of the results (cp3h_fetch_results)
    double* d_partials = nullptr;      // reduction scratch (cp3h_sqdiff_results_f32)
    double* d_like = nullptr;          // per-workgroup sums of the likelihood kernel
    unsigned like_blocks = 0;          // entries of d_like
    double* h_scalar = nullptr;        // pinned landing slot for a device scalar
    double* d_cos = nullptr;
    double* d_energies = nullptr;
    int* d_maxlayers = nullptr;
    double* d_radii = nullptr;
    double* d_rhos = nullptr;
    double* d_geom = nullptr;
    void* d_Q = nullptr;
    float* d_M = nullptr;              // mixed only: basis changes between shells
    double* d_R = nullptr;
    void* d_out = nullptr;
    int out_bytes = 8;
    std::vector<int> maxlayers;
    cp3::OscParams params{};
    bool has_energies = false, has_cosines = false, has_density = false, has_params = false;
    bool has_timing = false;
    bool timing = false;               // record HIP events around the kernels (cp3h_set_timing)

    std::size_t n_results() const { return std::size_t(9) * n_c * n_e; }
};

namespace {

int check_handle(cp3h_propagator* h) {
    if (!h) return fail(CP3H_ERR_ARGUMENT, "null propagator handle");
    return CP3H_OK;
}

// row mu of the mixed kernels' real basis, as the kernels read it
float4 wmu_f32(const cp3::OscParams& p) {
    return make_float4(float(p.Wmu[0]), float(p.Wmu[1]), float(p.Wmu[2]), float(p.Wmu[3]));
}

// the single-grid mixed kernel's vacuum constants for the handle's mixing matrix
cp3::VacuumBlock vacuum_of(const cp3::OscParams& p) {
    cp3::VacuumBlock b{};
    cp3host::vacuum_block(p.U, b.c);
    return b;
}

int sync_and_select(cp3h_propagator* h) {
    CP3H_HIP(hipSetDevice(h->device));
    CP3H_HIP(hipStreamSynchronize(h->stream));
    return CP3H_OK;
}

template <typename T>
void free_dev(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

int upload(cp3h_propagator* h, void* dst, const void* src, std::size_t bytes) {
    if (bytes == 0) return CP3H_OK;
    CP3H_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
    CP3H_HIP(hipStreamSynchronize(h->stream));
    return CP3H_OK;
}

}  // namespace

extern "C" {

const char* cp3h_last_error(void) { return cp3host::g_error.c_str(); }

int cp3h_device_count(int* count) {
    if (!count) return fail(CP3H_ERR_ARGUMENT, "null count");
    int n = 0;
    const hipError_t err = hipGetDeviceCount(&n);
    *count = err == hipSuccess ? n : 0;
    return CP3H_OK;
}

int cp3h_create(int device, int n_cosines, int n_energies, int precision, cp3h_propagator** out) {
    if (!out) return fail(CP3H_ERR_ARGUMENT, "null output handle pointer");
    *out = nullptr;
    if (n_cosines < 0 || n_energies < 0) return fail(CP3H_ERR_ARGUMENT, "negative grid size");
    if (precision != CP3H_FP64_PLAIN && precision != CP3H_FP64 && precision != CP3H_MIXED)
        return fail(CP3H_ERR_ARGUMENT, "unknown precision mode " + std::to_string(precision));
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev == 0) return fail(CP3H_ERR_DEVICE, "No GPU found");
    if (device < 0 || device >= n_dev)
        return fail(CP3H_ERR_DEVICE, "The requested GPU Id " + std::to_string(device) + " is not available.");

    cp3h_propagator* h = new cp3h_propagator();
    h->device = device;
    h->n_c = n_cosines;
    h->n_e = n_energies;
    h->precision = precision;
    h->ne_pad = int(cp3::round_up(unsigned(std::max(n_energies, 1)), cp3::kWave));
    h->out_bytes = precision == CP3H_MIXED ? 4 : 8;
    h->maxlayers.assign(n_cosines, 0);
    auto bail = [&](int code) { cp3h_destroy(h); return code; };
#define CP3H_TRY(call) do { int rc_ = (call); if (rc_ != CP3H_OK) return bail(rc_); } while (0)
    CP3H_TRY([&]() -> int {
        CP3H_HIP(hipSetDevice(device));
        CP3H_HIP(hipStreamCreate(&h->own_stream));
        h->stream = h->own_stream;
        CP3H_HIP(hipEventCreate(&h->ev_start));
        CP3H_HIP(hipEventCreate(&h->ev_pairs));
        CP3H_HIP(hipEventCreate(&h->ev_stop));
        CP3H_HIP(hipMalloc(&h->d_cos, sizeof(double) * std::max(n_cosines, 1)));
        CP3H_HIP(hipMalloc(&h->d_maxlayers, sizeof(int) * std::max(n_cosines, 1)));
        CP3H_HIP(hipMalloc(&h->d_energies, sizeof(double) * std::max(n_energies, 1)));
        CP3H_HIP(hipMalloc(&h->d_out, std::max<std::size_t>(h->n_results(), 1) * h->out_bytes));
        return CP3H_OK;
    }());
#undef CP3H_TRY
    *out = h;
    return CP3H_OK;
}

int cp3h_destroy(cp3h_propagator* h) {
    if (!h) return CP3H_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    free_dev(h->d_cos);
    free_dev(h->d_energies);
    free_dev(h->d_maxlayers);
    free_dev(h->d_radii);
    free_dev(h->d_rhos);
    free_dev(h->d_geom);
    free_dev(h->d_Q);
    free_dev(h->d_M);
    free_dev(h->d_R);
    free_dev(h->d_out);
    free_dev(h->d_partials);
    free_dev(h->d_like);
    if (h->h_scalar) (void)hipHostFree(h->h_scalar);
    if (h->h_results) (void)hipHostFree(h->h_results);
    if (h->ev_start) (void)hipEventDestroy(h->ev_start);
    if (h->ev_pairs) (void)hipEventDestroy(h->ev_pairs);
    if (h->ev_stop) (void)hipEventDestroy(h->ev_stop);
    if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
    delete h;
    return CP3H_OK;
}

int cp3h_set_energies_f64(cp3h_propagator* h, const double* energies, int n) {
    if (int rc = check_handle(h)) return rc;
    if (n != h->n_e)
        return fail(CP3H_ERR_ARGUMENT, "setEnergyList: propagator was not created for this number of energy nodes");
    if (n > 0 && !energies) return fail(CP3H_ERR_ARGUMENT, "null energy list");
    for (int i = 0; i < n; i++)
        if (!(energies[i] > 0.0)) return fail(CP3H_ERR_ARGUMENT, "setEnergyList: energies must be positive");
    if (int rc = sync_and_select(h)) return rc;
    if (int rc = upload(h, h->d_energies, energies, sizeof(double) * n)) return rc;
    h->has_energies = true;
    return CP3H_OK;
}

int cp3h_set_energies_f32(cp3h_propagator* h, const float* energies, int n) {
    if (int rc = check_handle(h)) return rc;
    if (n > 0 && !energies) return fail(CP3H_ERR_ARGUMENT, "null energy list");
    std::vector<double> e(energies, energies + std::max(n, 0));
    return cp3h_set_energies_f64(h, e.data(), n);
}

int cp3h_set_maxlayers(cp3h_propagator* h, const int* maxlayers, int n) {
    if (int rc = check_handle(h)) return rc;
    if (n != h->n_c) return fail(CP3H_ERR_ARGUMENT, "maxlayers size differs from the number of cosine nodes");
    if (n > 0 && !maxlayers) return fail(CP3H_ERR_ARGUMENT, "null maxlayers");
    for (int i = 0; i < n; i++)
        if (maxlayers[i] < 0) return fail(CP3H_ERR_ARGUMENT, "negative maxlayers entry");
    if (int rc = sync_and_select(h)) return rc;
    if (int rc = upload(h, h->d_maxlayers, maxlayers, sizeof(int) * n)) return rc;
    h->maxlayers.assign(maxlayers, maxlayers + n);
    return CP3H_OK;
}

int cp3h_set_cosines_f64(cp3h_propagator* h, const double* cosines, const int* maxlayers, int n) {
    if (int rc = check_handle(h)) return rc;
    if (n != h->n_c)
        return fail(CP3H_ERR_ARGUMENT, "setCosineList: propagator was not created for this number of cosine nodes");
    if (n > 0 && !cosines) return fail(CP3H_ERR_ARGUMENT, "null cosine list");
    for (int i = 0; i < n; i++)
        if (!(cosines[i] >= -1.0 && cosines[i] <= 1.0))
            return fail(CP3H_ERR_ARGUMENT, "setCosineList: cosines must lie in [-1, 1]");
    if (int rc = cp3h_set_maxlayers(h, maxlayers, n)) return rc;
    if (int rc = upload(h, h->d_cos, cosines, sizeof(double) * n)) return rc;
    h->has_cosines = true;
    return CP3H_OK;
}

int cp3h_set_cosines_f32(cp3h_propagator* h, const float* cosines, const int* maxlayers, int n) {
    if (int rc = check_handle(h)) return rc;
    if (n > 0 && !cosines) return fail(CP3H_ERR_ARGUMENT, "null cosine list");
    std::vector<double> c(cosines, cosines + std::max(n, 0));
    return cp3h_set_cosines_f64(h, c.data(), maxlayers, n);
}

int cp3h_set_density_f64(cp3h_propagator* h, const double* radii_desc, const double* rhos_desc, int n_rows) {
    if (int rc = check_handle(h)) return rc;
    if (n_rows <= 0) return fail(CP3H_ERR_ARGUMENT, "setDensity : vectors must not be empty");
    if (!radii_desc || !rhos_desc) return fail(CP3H_ERR_ARGUMENT, "null density arrays");
    for (int i = 1; i < n_rows; i++)
        if (radii_desc[i] > radii_desc[i - 1])
            return fail(CP3H_ERR_ARGUMENT, "setDensity: radii must be in descending order");
    if (int rc = sync_and_select(h)) return rc;
    if (n_rows != h->rows) {
        free_dev(h->d_radii);
        free_dev(h->d_rhos);
        free_dev(h->d_geom);
        free_dev(h->d_Q);
        free_dev(h->d_M);
        free_dev(h->d_R);
        h->rows = 0;
        h->has_density = false;
        CP3H_HIP(hipMalloc(&h->d_radii, sizeof(double) * n_rows));
        CP3H_HIP(hipMalloc(&h->d_rhos, sizeof(double) * n_rows));
        const std::size_t slots = std::size_t(n_rows) + 1;
        CP3H_HIP(hipMalloc(&h->d_geom, sizeof(double) * std::max<std::size_t>(std::size_t(h->n_c) * slots, 1)));
        if (h->precision != CP3H_FP64_PLAIN) {
            const std::size_t cells = slots * h->ne_pad;
            const bool mixed = h->precision == CP3H_MIXED;
            CP3H_HIP(hipMalloc(&h->d_Q, cells * (mixed ? cp3::kMixedQFloats * sizeof(float) : 27 * sizeof(double))));
            if (mixed) CP3H_HIP(hipMalloc(&h->d_M, cells * cp3::kMixedMFloats * sizeof(float)));
            CP3H_HIP(hipMalloc(&h->d_R, cells * (mixed ? cp3::kMixedRDoubles : 3) * sizeof(double)));
        }
        h->rows = n_rows;
    }
    if (int rc = upload(h, h->d_radii, radii_desc, sizeof(double) * n_rows)) return rc;
    if (int rc = upload(h, h->d_rhos, rhos_desc, sizeof(double) * n_rows)) return rc;
    h->has_density = true;
    return CP3H_OK;
}

int cp3h_set_density_f32(cp3h_propagator* h, const float* radii_desc, const float* rhos_desc, int n_rows) {
    if (int rc = check_handle(h)) return rc;
    if (n_rows > 0 && (!radii_desc || !rhos_desc)) return fail(CP3H_ERR_ARGUMENT, "null density arrays");
    std::vector<double> r(radii_desc, radii_desc + std::max(n_rows, 0));
    std::vector<double> d(rhos_desc, rhos_desc + std::max(n_rows, 0));
    return cp3h_set_density_f64(h, r.data(), d.data(), n_rows);
}

int cp3h_set_params_f64(cp3h_propagator* h, const double* U18, const double* DM9) {
    if (int rc = check_handle(h)) return rc;
    if (!U18 || !DM9) return fail(CP3H_ERR_ARGUMENT, "null parameter arrays");
    std::memcpy(h->params.U, U18, sizeof(double) * 18);
    std::memcpy(h->params.DM, DM9, sizeof(double) * 9);
    vacuum_order<double>(DM9, h->params.order);
    cp3host::real_basis(h->params.U, h->params.DM, h->params.Hr, h->params.Wmu);
    h->has_params = true;
    return CP3H_OK;
}

int cp3h_set_params_f32(cp3h_propagator* h, const float* U18, const float* DM9) {
    if (int rc = check_handle(h)) return rc;
    if (!U18 || !DM9) return fail(CP3H_ERR_ARGUMENT, "null parameter arrays");
    for (int i = 0; i < 18; i++) h->params.U[i] = U18[i];
    for (int i = 0; i < 9; i++) h->params.DM[i] = DM9[i];
    vacuum_order<float>(DM9, h->params.order);
    cp3host::real_basis(h->params.U, h->params.DM, h->params.Hr, h->params.Wmu);
    h->has_params = true;
    return CP3H_OK;
}

}  // extern "C"

namespace {

// Checks that a computation can be launched and returns its grid description.
int ready_to_compute(cp3h_propagator* h, int antineutrino, double production_height_cm, cp3::GridDesc& g) {
    if (int rc = check_handle(h)) return rc;
    if (antineutrino != 0 && antineutrino != 1) return fail(CP3H_ERR_ARGUMENT, "neutrino type must be 0 or 1");
    if (!h->has_energies) return fail(CP3H_ERR_STATE, "calculateProbabilities: energy list was not set");
    if (!h->has_cosines) return fail(CP3H_ERR_STATE, "calculateProbabilities: cosine list was not set");
    if (!h->has_density) return fail(CP3H_ERR_STATE, "calculateProbabilities: density was not set");
    if (!h->has_params) return fail(CP3H_ERR_STATE, "calculateProbabilities: oscillation parameters were not set");
    if (!(production_height_cm >= 0.0)) return fail(CP3H_ERR_ARGUMENT, "production height must be >= 0");
    for (int i = 0; i < h->n_c; i++)
        if (h->maxlayers[i] > h->rows)
            return fail(CP3H_ERR_STATE, "maxlayers exceeds the number of shells of the density model");
    CP3H_HIP(hipSetDevice(h->device));
    h->params.antinu = antineutrino;
    g = cp3::GridDesc{h->d_cos, h->d_energies, h->d_maxlayers, h->d_radii, h->d_rhos,
                      h->n_c, h->n_e, h->rows, production_height_cm};
    return CP3H_OK;
}

// The set-up kernels of a computation (none for CP3H_FP64_PLAIN): layer lengths and the
// per-(shell, energy) spectrum table.
int enqueue_setup(cp3h_propagator* h, const cp3::GridDesc& g) {
    if (h->n_c == 0 || h->n_e == 0 || h->precision == CP3H_FP64_PLAIN) return CP3H_OK;
    if (h->precision == CP3H_MIXED) {
        CP3H_HIP(cp3::launch_setup_mixed(g, h->params, h->ne_pad, h->d_geom, static_cast<float*>(h->d_Q), h->d_M,
                                         h->d_R, h->stream));
    } else {
        CP3H_HIP(cp3::launch_geometry(g, h->d_geom, h->stream));
        CP3H_HIP(cp3::launch_spectrum_f64(g, h->params, h->ne_pad, static_cast<double*>(h->d_Q), h->d_R, h->stream));
    }
    return CP3H_OK;
}

// The per-pair kernel of a computation; with data != nullptr the likelihood variant (mixed only),
// whose per-workgroup sums go to d_like.
int enqueue_pairs(cp3h_propagator* h, const cp3::GridDesc& g, const float* data) {
    if (h->n_c == 0 || h->n_e == 0) return CP3H_OK;
    if (h->precision == CP3H_FP64_PLAIN) {
        CP3H_HIP(cp3::launch_plain_f64(g, h->params, static_cast<double*>(h->d_out), h->stream));
    } else if (h->precision == CP3H_MIXED) {
        if (data)
            CP3H_HIP(cp3::launch_pairs_mixed_sqdiff(g, h->ne_pad, h->d_geom, static_cast<const float*>(h->d_Q), h->d_M,
                                                    h->d_R, wmu_f32(h->params), vacuum_of(h->params), data, h->d_like,
                                                    h->stream));
        else
            CP3H_HIP(cp3::launch_pairs_mixed(g, h->ne_pad, h->d_geom, static_cast<const float*>(h->d_Q), h->d_M,
                                             h->d_R, wmu_f32(h->params), vacuum_of(h->params),
                                             static_cast<float*>(h->d_out), h->stream));
    } else {
        CP3H_HIP(cp3::launch_pairs_f64(g, h->ne_pad, h->d_geom, static_cast<const double*>(h->d_Q), h->d_R,
                                       static_cast<double*>(h->d_out), h->stream));
    }
    return CP3H_OK;
}

// Scratch of the likelihood kernels (allocated on first use).
int ensure_likelihood_scratch(cp3h_propagator* h, unsigned blocks) {
    if (!h->d_partials) {
        CP3H_HIP(hipMalloc(&h->d_partials, sizeof(double) * (cp3::reduce_partials_needed() + 1)));
        CP3H_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->h_scalar), sizeof(double)));
    }
    if (blocks > h->like_blocks) {
        free_dev(h->d_like);
        CP3H_HIP(hipMalloc(&h->d_like, sizeof(double) * blocks));
        h->like_blocks = blocks;
    }
    return CP3H_OK;
}

}  // namespace

extern "C" {

int cp3h_compute_async(cp3h_propagator* h, int antineutrino, double production_height_cm) {
    cp3::GridDesc g;
    if (int rc = ready_to_compute(h, antineutrino, production_height_cm, g)) return rc;
    if (h->timing) CP3H_HIP(hipEventRecord(h->ev_start, h->stream));
    if (int rc = enqueue_setup(h, g)) return rc;
    if (h->timing) CP3H_HIP(hipEventRecord(h->ev_pairs, h->stream));
    if (int rc = enqueue_pairs(h, g, nullptr)) return rc;
    if (h->timing) CP3H_HIP(hipEventRecord(h->ev_stop, h->stream));
    h->has_timing = h->timing;
    return CP3H_OK;
}

int cp3h_time_kernels(cp3h_propagator* h, int antineutrino, double production_height_cm, const float* data_device,
                      int reps, float* setup_ms, float* pairs_ms) {
    cp3::GridDesc g;
    if (int rc = ready_to_compute(h, antineutrino, production_height_cm, g)) return rc;
    if (reps < 1) return fail(CP3H_ERR_ARGUMENT, "reps must be >= 1");
    if (data_device) {
        if (h->precision != CP3H_MIXED) return fail(CP3H_ERR_STATE, "the likelihood kernel needs a CP3H_MIXED handle");
        if (int rc = ensure_likelihood_scratch(h, cp3::likelihood_blocks(g))) return rc;
    }
    float ms = 0.0f;
    CP3H_HIP(hipEventRecord(h->ev_start, h->stream));
    for (int r = 0; r < reps; r++)
        if (int rc = enqueue_setup(h, g)) return rc;
    CP3H_HIP(hipEventRecord(h->ev_pairs, h->stream));
    for (int r = 0; r < reps; r++)
        if (int rc = enqueue_pairs(h, g, data_device)) return rc;
    CP3H_HIP(hipEventRecord(h->ev_stop, h->stream));
    CP3H_HIP(hipEventSynchronize(h->ev_stop));
    CP3H_HIP(hipEventElapsedTime(&ms, h->ev_start, h->ev_pairs));
    if (setup_ms) *setup_ms = ms / reps;
    CP3H_HIP(hipEventElapsedTime(&ms, h->ev_pairs, h->ev_stop));
    if (pairs_ms) *pairs_ms = ms / reps;
    h->has_timing = false;       // the events no longer bracket one computation
    return CP3H_OK;
}

int cp3h_wait(cp3h_propagator* h) {
    if (int rc = check_handle(h)) return rc;
    return sync_and_select(h);
}

int cp3h_compute(cp3h_propagator* h, int antineutrino, double production_height_cm) {
    if (int rc = cp3h_compute_async(h, antineutrino, production_height_cm)) return rc;
    return cp3h_wait(h);
}

int cp3h_fetch_results(cp3h_propagator* h, const void** host_ptr, int* elem_bytes) {
    if (int rc = check_handle(h)) return rc;
    if (!host_ptr || !elem_bytes) return fail(CP3H_ERR_ARGUMENT, "null output pointer");
    CP3H_HIP(hipSetDevice(h->device));
    const std::size_t bytes = std::max<std::size_t>(h->n_results(), 1) * h->out_bytes;
    if (!h->h_results) CP3H_HIP(hipHostMalloc(&h->h_results, bytes));
    if (h->n_results() > 0)
        CP3H_HIP(hipMemcpyAsync(h->h_results, h->d_out, h->n_results() * h->out_bytes, hipMemcpyDeviceToHost,
                                h->stream));
    CP3H_HIP(hipStreamSynchronize(h->stream));
    *host_ptr = h->h_results;
    *elem_bytes = h->out_bytes;
    return CP3H_OK;
}

}  // extern "C"

namespace {

// D2H into the pinned copy, then a host copy / conversion into the caller's array.
template <typename T>
int copy_results(cp3h_propagator* h, T* dst, std::size_t count) {
    if (int rc = check_handle(h)) return rc;
    if (count != h->n_results()) return fail(CP3H_ERR_ARGUMENT, "result count must be 9*n_cosines*n_energies");
    if (count > 0 && !dst) return fail(CP3H_ERR_ARGUMENT, "null destination");
    const void* src = nullptr;
    int bytes = 0;
    if (int rc = cp3h_fetch_results(h, &src, &bytes)) return rc;
    if (bytes == 8) {
        const double* s = static_cast<const double*>(src);
        std::copy(s, s + count, dst);
    } else {
        const float* s = static_cast<const float*>(src);
        std::copy(s, s + count, dst);
    }
    return CP3H_OK;
}

}  // namespace

extern "C" {

int cp3h_copy_results_f64(cp3h_propagator* h, double* dst, size_t count) { return copy_results(h, dst, count); }

int cp3h_copy_results_f32(cp3h_propagator* h, float* dst, size_t count) { return copy_results(h, dst, count); }

int cp3h_device_results(cp3h_propagator* h, const void** ptr, int* elem_bytes) {
    if (int rc = check_handle(h)) return rc;
    if (!ptr || !elem_bytes) return fail(CP3H_ERR_ARGUMENT, "null output pointer");
    *ptr = h->d_out;
    *elem_bytes = h->out_bytes;
    return CP3H_OK;
}

int cp3h_stream(cp3h_propagator* h, void** stream) {
    if (int rc = check_handle(h)) return rc;
    if (!stream) return fail(CP3H_ERR_ARGUMENT, "null output pointer");
    *stream = static_cast<void*>(h->stream);
    return CP3H_OK;
}

int cp3h_set_stream(cp3h_propagator* h, void* stream) {
    if (int rc = check_handle(h)) return rc;
    if (int rc = sync_and_select(h)) return rc;
    h->stream = stream ? static_cast<hipStream_t>(stream) : h->own_stream;
    return CP3H_OK;
}

int cp3h_copy_results_device(cp3h_propagator* h, void* dst_device, size_t count) {
    if (int rc = check_handle(h)) return rc;
    if (count != h->n_results()) return fail(CP3H_ERR_ARGUMENT, "result count must be 9*n_cosines*n_energies");
    if (count > 0 && !dst_device) return fail(CP3H_ERR_ARGUMENT, "null destination");
    CP3H_HIP(hipSetDevice(h->device));
    if (count) CP3H_HIP(hipMemcpyAsync(dst_device, h->d_out, count * h->out_bytes, hipMemcpyDeviceToDevice, h->stream));
    CP3H_HIP(hipStreamSynchronize(h->stream));
    return CP3H_OK;
}

int cp3h_sqdiff_results_f32(cp3h_propagator* h, const float* data_device, double* result) {
    if (int rc = check_handle(h)) return rc;
    if (h->out_bytes != 4) return fail(CP3H_ERR_STATE, "sqdiff_results_f32 needs float results (CP3H_MIXED)");
    if (!data_device || !result) return fail(CP3H_ERR_ARGUMENT, "null pointer");
    if (reinterpret_cast<std::uintptr_t>(data_device) % 16 != 0)
        return fail(CP3H_ERR_ARGUMENT, "data must be 16-byte aligned");
    CP3H_HIP(hipSetDevice(h->device));
    if (!h->d_partials) {
        CP3H_HIP(hipMalloc(&h->d_partials, sizeof(double) * (cp3::reduce_partials_needed() + 1)));
        CP3H_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->h_scalar), sizeof(double)));
    }
    CP3H_HIP(cp3::launch_sqdiff_sum(static_cast<const float*>(h->d_out), data_device, h->n_results(), h->d_partials,
                                    h->stream));
    CP3H_HIP(hipMemcpyAsync(h->h_scalar, h->d_partials + cp3::reduce_partials_needed(), sizeof(double),
                            hipMemcpyDeviceToHost, h->stream));
    CP3H_HIP(hipStreamSynchronize(h->stream));
    *result = *h->h_scalar;
    return CP3H_OK;
}

int cp3h_compute_sqdiff_f32(cp3h_propagator* h, int antineutrino, double production_height_cm,
                            const float* data_device, double* result) {
    if (int rc = check_handle(h)) return rc;
    if (h->precision != CP3H_MIXED) return fail(CP3H_ERR_STATE, "compute_sqdiff_f32 needs a CP3H_MIXED handle");
    if (!result || (h->n_results() > 0 && !data_device)) return fail(CP3H_ERR_ARGUMENT, "null pointer");
    cp3::GridDesc g;
    if (int rc = ready_to_compute(h, antineutrino, production_height_cm, g)) return rc;
    const unsigned blocks = cp3::likelihood_blocks(g);
    if (blocks == 0) {
        *result = 0.0;
        return CP3H_OK;
    }
    if (int rc = ensure_likelihood_scratch(h, blocks)) return rc;
    if (h->timing) CP3H_HIP(hipEventRecord(h->ev_start, h->stream));
    if (int rc = enqueue_setup(h, g)) return rc;
    if (h->timing) CP3H_HIP(hipEventRecord(h->ev_pairs, h->stream));
    if (int rc = enqueue_pairs(h, g, data_device)) return rc;
    if (h->timing) CP3H_HIP(hipEventRecord(h->ev_stop, h->stream));
    h->has_timing = h->timing;
    CP3H_HIP(cp3::launch_sum_f64(h->d_like, blocks, h->d_partials, h->stream));
    CP3H_HIP(hipMemcpyAsync(h->h_scalar, h->d_partials + cp3::reduce_partials_needed(), sizeof(double),
                            hipMemcpyDeviceToHost, h->stream));
    CP3H_HIP(hipStreamSynchronize(h->stream));
    *result = *h->h_scalar;
    return CP3H_OK;
}

int cp3h_set_timing(cp3h_propagator* h, int enable) {
    if (int rc = check_handle(h)) return rc;
    h->timing = enable != 0;
    return CP3H_OK;
}

int cp3h_last_compute_ms(cp3h_propagator* h, float* total_ms, float* pairs_ms) {
    if (int rc = check_handle(h)) return rc;
    if (!h->has_timing)
        return fail(CP3H_ERR_STATE, "no compute has been launched with timing enabled (cp3h_set_timing)");
    CP3H_HIP(hipSetDevice(h->device));
    CP3H_HIP(hipEventSynchronize(h->ev_stop));
    if (total_ms) CP3H_HIP(hipEventElapsedTime(total_ms, h->ev_start, h->ev_stop));
    if (pairs_ms) CP3H_HIP(hipEventElapsedTime(pairs_ms, h->ev_pairs, h->ev_stop));
    return CP3H_OK;
}

}  // extern "C"
