// Allocation-free mixed-precision propagation: cp3h_propagate_mixed_noalloc (include/cp3hip.h).
//
// The same set-up and pair kernels as cp3h_compute_async on a CP3H_MIXED handle, with every
// input and the output in caller memory and the kernels' scratch (per-cosine layer lengths and
// the per-(shell, energy) spectrum table) in one file-scope device array instead of handle
// buffers.  No hipMalloc / hipFree / host allocation happens here, so the call can sit in code
// that must not allocate (HIP graph capture, serving loops).
//
// Scratch: kScratchBytes per GPU, held from the library's load until the process ends.  Its
// device address is looked up once per device (hipGetSymbolAddress, with that device current)
// on the first call that runs on it and cached; host code only ever passes that address to the
// kernels.  Every call on one GPU shares that GPU's array: two calls whose work overlaps on the
// device (different streams, no ordering between them) race.  Calls on one stream are ordered.
// Every scratch cell a computation reads is written by its own set-up kernel first, so nothing
// has to be cleared between calls.
#include "../../include/cp3hip.h"
#include "common.hpp"
#include "host_util.hpp"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>

namespace {
constexpr std::size_t kScratchBytes = std::size_t(128) << 20;   // 128 MiB per GPU
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

using cp3host::fail;

std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (the current device), looked up on first use.
int scratch_address(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !ptr)
            return fail(CP3H_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(err));
        g_scratch_ptr[device] = ptr;
    }
    *out = static_cast<unsigned char*>(g_scratch_ptr[device]);
    return CP3H_OK;
}

std::size_t align256(std::size_t x) { return (x + 255) / 256 * 256; }

}  // namespace

extern "C" int cp3h_propagate_mixed_noalloc(int device, const double* cosines_device, const int* maxlayers_device,
                                            int n_cosines, const double* energies_device, int n_energies,
                                            const double* radii_desc_device, const double* rhos_desc_device,
                                            int n_rows, const double* U18, const double* DM9, int antineutrino,
                                            double production_height_cm, float* out_device, void* stream) {
    if (n_cosines < 0 || n_energies < 0) return fail(CP3H_ERR_ARGUMENT, "negative grid size");
    if (n_rows <= 0) return fail(CP3H_ERR_ARGUMENT, "setDensity : vectors must not be empty");
    if (!U18 || !DM9) return fail(CP3H_ERR_ARGUMENT, "null parameter arrays");
    if (antineutrino != 0 && antineutrino != 1) return fail(CP3H_ERR_ARGUMENT, "neutrino type must be 0 or 1");
    if (!(production_height_cm >= 0.0)) return fail(CP3H_ERR_ARGUMENT, "production height must be >= 0");
    if (!radii_desc_device || !rhos_desc_device) return fail(CP3H_ERR_ARGUMENT, "null density arrays");
    const std::size_t pairs = std::size_t(n_cosines) * std::size_t(n_energies);
    if (pairs > 0 && (!cosines_device || !maxlayers_device || !energies_device || !out_device))
        return fail(CP3H_ERR_ARGUMENT, "null grid or result array");
    // scratch layout: layer lengths, then the two spectrum-table arrays, each 256-byte aligned
    // (every factor is first bounded by the limit, so the products below cannot wrap)
    if (std::size_t(n_cosines) > kScratchBytes || std::size_t(n_rows) >= kScratchBytes ||
        std::size_t(n_energies) > kScratchBytes)
        return fail(CP3H_ERR_ARGUMENT, "grid too large for the allocation-free scratch (limit " +
                                           std::to_string(kScratchBytes) + " bytes)");
    const std::size_t slots = std::size_t(n_rows) + 1;
    const std::size_t ne_pad = cp3::round_up(unsigned(n_energies > 0 ? n_energies : 1), cp3::kWave);
    const std::size_t geom_bytes = align256(sizeof(double) * std::size_t(n_cosines) * slots);
    const std::size_t q_bytes = align256(sizeof(float) * cp3::kMixedQFloats * slots * ne_pad);
    const std::size_t m_bytes = align256(sizeof(float) * cp3::kMixedMFloats * slots * ne_pad);
    const std::size_t r_bytes = align256(sizeof(double) * cp3::kMixedRDoubles * slots * ne_pad);
    const std::size_t need = geom_bytes + q_bytes + m_bytes + r_bytes;
    if (need > kScratchBytes)
        return fail(CP3H_ERR_ARGUMENT, "grid too large for the allocation-free scratch: needs " +
                                           std::to_string(need) + " bytes, the limit is " +
                                           std::to_string(kScratchBytes) +
                                           " (8 n_cosines (n_rows+1) + 96 (n_rows+1) n_energies rounded up to 64)");

    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev == 0) return fail(CP3H_ERR_DEVICE, "No GPU found");
    if (device < 0 || device >= n_dev || device >= kMaxDevices)
        return fail(CP3H_ERR_DEVICE, "The requested GPU Id " + std::to_string(device) + " is not available.");

    CP3H_HIP(hipSetDevice(device));
    unsigned char* base = nullptr;
    if (int rc = scratch_address(device, &base)) return rc;
    if (pairs == 0) return CP3H_OK;

    cp3::OscParams p{};
    for (int i = 0; i < 18; i++) p.U[i] = U18[i];
    for (int i = 0; i < 9; i++) p.DM[i] = DM9[i];
    cp3host::vacuum_order<double>(DM9, p.order);
    cp3host::real_basis(p.U, p.DM, p.Hr, p.Wmu);
    p.antinu = antineutrino;
    const float4 wmu = make_float4(float(p.Wmu[0]), float(p.Wmu[1]), float(p.Wmu[2]), float(p.Wmu[3]));
    cp3::VacuumBlock vac{};
    cp3host::vacuum_block(p.U, vac.c);

    double* geom = reinterpret_cast<double*>(base);
    float* Q = reinterpret_cast<float*>(base + geom_bytes);
    float* M = reinterpret_cast<float*>(base + geom_bytes + q_bytes);
    double* R = reinterpret_cast<double*>(base + geom_bytes + q_bytes + m_bytes);
    const cp3::GridDesc g{cosines_device, energies_device, maxlayers_device, radii_desc_device, rhos_desc_device,
                          n_cosines, n_energies, n_rows, production_height_cm};
    hipStream_t s = static_cast<hipStream_t>(stream);
    CP3H_HIP(cp3::launch_setup_mixed(g, p, int(ne_pad), geom, Q, M, R, s));
    CP3H_HIP(cp3::launch_pairs_mixed(g, int(ne_pad), geom, Q, M, R, wmu, vac, out_device, s));
    return CP3H_OK;
}
