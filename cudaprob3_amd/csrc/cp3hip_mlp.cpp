// Host side of the surrogate network API (cp3h_mlp_* in include/cp3hip.h): parameters, Adam
// state and activation buffers on the device, and the launch sequences of inference, the
// backward pass and the optimizer step.  All launches go to one stream.
#include "../../include/cp3hip.h"
#include "common.hpp"
#include "host_util.hpp"
#include "mlp.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

using cp3host::fail;

namespace {

struct LayerShape {
    int N, K;            // padded
    int n_valid, k_valid;
};

constexpr LayerShape kShape[cp3::kMlpLayers] = {
    {cp3::kMlpHidden, cp3::kMlpInPad, cp3::kMlpHidden, cp3::kMlpIn},
    {cp3::kMlpHidden, cp3::kMlpHidden, cp3::kMlpHidden, cp3::kMlpHidden},
    {cp3::kMlpHidden, cp3::kMlpHidden, cp3::kMlpHidden, cp3::kMlpHidden},
    {cp3::kMlpHidden, cp3::kMlpHidden, cp3::kMlpHidden, cp3::kMlpHidden},
    {cp3::kMlpHidden, cp3::kMlpHidden, cp3::kMlpHidden, cp3::kMlpHidden},
    {cp3::kMlpOutPad, cp3::kMlpHidden, cp3::kMlpOut, cp3::kMlpHidden},
};

constexpr int kSplit = 8;   // split-K parts of the weight-gradient GEMMs

// bf16 forward weights of layer l: [N][wb_ld(l)], each row's K values at 0 and at wb_dup(l)
constexpr int wb_ld(int l) { return l == 0 ? cp3::kMlpInPad : 2 * kShape[l].K; }
constexpr int wb_dup(int l) { return l == 0 ? cp3::kMlpInLo : kShape[l].K; }

std::size_t param_count() {
    std::size_t n = 0;
    for (const auto& s : kShape) n += std::size_t(s.n_valid) * s.k_valid + s.n_valid;
    return n;
}

}  // namespace

struct cp3h_mlp {
    int device = 0, max_batch = 0, cu_count = 1;
    bool training = false;
    hipStream_t stream = nullptr, own_stream = nullptr;
    float* w[cp3::kMlpLayers] = {};
    float* b[cp3::kMlpLayers] = {};
    __bf16* wb[cp3::kMlpLayers] = {};
    __bf16* wtb[cp3::kMlpLayers] = {};
    float *m1w[cp3::kMlpLayers] = {}, *m2w[cp3::kMlpLayers] = {}, *m1b[cp3::kMlpLayers] = {},
          *m2b[cp3::kMlpLayers] = {};
    float* gw[cp3::kMlpLayers] = {};
    float* gb[cp3::kMlpLayers] = {};
    float* partials = nullptr;
    __bf16 *x0 = nullptr, *x0t = nullptr;
    float* z[cp3::kMlpHiddenLayers] = {};
    __bf16* y[cp3::kMlpHiddenLayers] = {};
    __bf16* yt[cp3::kMlpHiddenLayers] = {};
    float* zout = nullptr;
    __bf16 *dz[2] = {}, *dzt[2] = {};
    float* loss_part = nullptr;
    float* loss_dev = nullptr;
    float* loss_host = nullptr;
    cp3::MlpScaling scaling{};
    int adam_t = 0;
    int forward_kernel = CP3H_MLP_FUSED4;
    std::vector<void*> allocations;
};

namespace {

template <typename T>
int dev_alloc(cp3h_mlp* h, T** p, std::size_t count) {
    CP3H_HIP(hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * std::max<std::size_t>(count, 1)));
    h->allocations.push_back(*p);
    CP3H_HIP(hipMemsetAsync(*p, 0, sizeof(T) * std::max<std::size_t>(count, 1), h->stream));
    return CP3H_OK;
}

int check(cp3h_mlp* h) {
    if (!h) return fail(CP3H_ERR_ARGUMENT, "null network handle");
    return CP3H_OK;
}

// write the bf16 copies of the fp32 weights (an Adam call that updates nothing)
int refresh_bf16(cp3h_mlp* h) {
    const cp3::AdamStep none{0.0f, 0.0f, 0.0f, 1.0f, 1.0f, 1.0f};
    for (int l = 0; l < cp3::kMlpLayers; l++)
        CP3H_HIP(cp3::launch_adam_layer(h->w[l], nullptr, nullptr, nullptr, kShape[l].N, kShape[l].K, 0, 0, none,
                                        h->wb[l], wb_ld(l), wb_dup(l), h->training ? h->wtb[l] : nullptr, h->stream));
    return CP3H_OK;
}

// forward pass of M rows of x (device fp32 [M][9]); training keeps Z, Y^T and the transposed
// input for the backward pass.  With targets, also the loss partials and dZ of the output.
int forward(cp3h_mlp* h, const float* x, int M, const float* targets, float* y_out) {
    const bool train = targets != nullptr;
    CP3H_HIP(cp3::launch_featurize(x, M, h->scaling, h->x0, train ? h->x0t : nullptr, h->stream));
    const __bf16* prev = h->x0;
    for (int l = 0; l < cp3::kMlpHiddenLayers; l++) {
        const LayerShape& s = kShape[l];
        const int kk = wb_ld(l);   // hi and lo terms of the previous activations
        // inference ping-pongs between two activation buffers
        __bf16* out = train ? h->y[l] : h->y[l % 2];
        cp3::GemmDesc d{prev, kk, h->wb[l], kk, M, s.N, kk, 1, cp3::kEpiBiasGelu, h->b[l], nullptr, 0,
                        train ? h->z[l] : nullptr, s.N, out, cp3::kMlpActLd, train ? h->yt[l] : nullptr, M,
                        s.n_valid, cp3::kMlpHidden};
        CP3H_HIP(cp3::launch_gemm_nt(d, h->stream));
        prev = out;
    }
    constexpr int lo = cp3::kMlpLayers - 1;
    const LayerShape& o = kShape[lo];
    cp3::GemmDesc d{prev, wb_ld(lo), h->wb[lo], wb_ld(lo), M, o.N, wb_ld(lo), 1, cp3::kEpiBias, h->b[lo], nullptr,
                    0, h->zout, o.N, nullptr, 0, nullptr, 0, o.n_valid, 0};
    CP3H_HIP(cp3::launch_gemm_nt(d, h->stream));
    CP3H_HIP(cp3::launch_output_head(h->zout, M, targets, y_out, train ? h->dz[0] : nullptr,
                                     train ? h->dzt[0] : nullptr, train ? h->loss_part : nullptr, h->stream));
    if (train) CP3H_HIP(cp3::launch_sum_floats(h->loss_part, cp3::output_head_blocks(M), h->loss_dev, h->stream));
    return CP3H_OK;
}

// backward pass after forward(train): gradients into gw, gb
int backward(cp3h_mlp* h, int M) {
    int cur = 0;   // dz[cur] / dzt[cur] hold dZ of layer l ([M][N_l], [N_l][M])
    for (int l = cp3::kMlpLayers - 1; l >= 0; l--) {
        const LayerShape& s = kShape[l];
        const __bf16* prev_t = l == 0 ? h->x0t : h->yt[l - 1];
        // dW = dZ^T Y_{l-1}: A = dZ^T [N][M], B = Y_{l-1}^T [K][M], reduction over M in kSplit parts
        cp3::GemmDesc gwd{h->dzt[cur], M, prev_t, M, s.N, s.K, M, kSplit, cp3::kEpiPartial, nullptr, nullptr, 0,
                          h->partials, s.K, nullptr, 0, nullptr, 0, s.N, 0};
        CP3H_HIP(cp3::launch_gemm_nt(gwd, h->stream));
        CP3H_HIP(cp3::launch_sum_parts(h->partials, kSplit, s.N * s.K, h->gw[l], h->stream));
        CP3H_HIP(cp3::launch_row_sums(h->dzt[cur], s.N, M, h->gb[l], h->stream));
        if (l > 0) {
            // dY_{l-1} = dZ W: A = dZ [M][N], B = W^T [K][N]; epilogue applies gelu'(Z_{l-1})
            cp3::GemmDesc gx{h->dz[cur], s.N, h->wtb[l], s.N, M, s.K, s.N, 1, cp3::kEpiGeluGrad, nullptr,
                             h->z[l - 1], s.K, nullptr, 0, h->dz[1 - cur], s.K, h->dzt[1 - cur], M, s.k_valid, 0};
            CP3H_HIP(cp3::launch_gemm_nt(gx, h->stream));
            cur = 1 - cur;
        }
    }
    return CP3H_OK;
}

int loss_to_host(cp3h_mlp* h, int M, float* loss) {
    CP3H_HIP(hipMemcpyAsync(h->loss_host, h->loss_dev, sizeof(float), hipMemcpyDeviceToHost, h->stream));
    CP3H_HIP(hipStreamSynchronize(h->stream));
    if (loss) *loss = *h->loss_host / (float(cp3::kMlpOut) * float(M));
    return CP3H_OK;
}

int check_batch(cp3h_mlp* h, int M) {
    if (!h->training) return fail(CP3H_ERR_STATE, "network was created for inference only");
    if (M <= 0 || M > h->max_batch || M % (32 * kSplit) != 0)
        return fail(CP3H_ERR_ARGUMENT, "training batch must be a positive multiple of 256 up to max_batch");
    return CP3H_OK;
}

}  // namespace

extern "C" {

int cp3h_mlp_param_count(size_t* n) {
    if (!n) return fail(CP3H_ERR_ARGUMENT, "null output");
    *n = param_count();
    return CP3H_OK;
}

int cp3h_mlp_create(int device, int max_batch, int training, cp3h_mlp** out) {
    if (!out) return fail(CP3H_ERR_ARGUMENT, "null output handle pointer");
    *out = nullptr;
    if (max_batch <= 0 || max_batch % 256 != 0) return fail(CP3H_ERR_ARGUMENT, "max_batch must be a positive multiple of 256");
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev == 0) return fail(CP3H_ERR_DEVICE, "No GPU found");
    if (device < 0 || device >= n_dev)
        return fail(CP3H_ERR_DEVICE, "The requested GPU Id " + std::to_string(device) + " is not available.");
    cp3h_mlp* h = new cp3h_mlp();
    h->device = device;
    h->max_batch = max_batch;
    h->training = training != 0;
    for (int c = 0; c < cp3::kMlpIn; c++) h->scaling.scale[c] = 1.0f;
    const int rc = [&]() -> int {
        CP3H_HIP(hipSetDevice(device));
        CP3H_HIP(hipDeviceGetAttribute(&h->cu_count, hipDeviceAttributeMultiprocessorCount, device));
        CP3H_HIP(hipStreamCreate(&h->own_stream));
        h->stream = h->own_stream;
        const std::size_t M = max_batch;
        for (int l = 0; l < cp3::kMlpLayers; l++) {
            const std::size_t nk = std::size_t(kShape[l].N) * kShape[l].K;
            if (int e = dev_alloc(h, &h->w[l], nk)) return e;
            if (int e = dev_alloc(h, &h->b[l], kShape[l].N)) return e;
            if (int e = dev_alloc(h, &h->wb[l], std::size_t(kShape[l].N) * wb_ld(l))) return e;
            if (h->training) {
                if (int e = dev_alloc(h, &h->wtb[l], nk)) return e;
                if (int e = dev_alloc(h, &h->m1w[l], nk)) return e;
                if (int e = dev_alloc(h, &h->m2w[l], nk)) return e;
                if (int e = dev_alloc(h, &h->m1b[l], kShape[l].N)) return e;
                if (int e = dev_alloc(h, &h->m2b[l], kShape[l].N)) return e;
                if (int e = dev_alloc(h, &h->gw[l], nk)) return e;
                if (int e = dev_alloc(h, &h->gb[l], kShape[l].N)) return e;
            }
        }
        if (int e = dev_alloc(h, &h->x0, M * cp3::kMlpInPad)) return e;
        if (int e = dev_alloc(h, &h->zout, M * cp3::kMlpOutPad)) return e;
        if (h->training) {
            if (int e = dev_alloc(h, &h->partials, std::size_t(kSplit) * cp3::kMlpHidden * cp3::kMlpHidden)) return e;
            if (int e = dev_alloc(h, &h->x0t, M * cp3::kMlpInPad)) return e;
            for (int l = 0; l < cp3::kMlpHiddenLayers; l++) {
                if (int e = dev_alloc(h, &h->z[l], M * cp3::kMlpHidden)) return e;
                if (int e = dev_alloc(h, &h->y[l], M * cp3::kMlpActLd)) return e;
                if (int e = dev_alloc(h, &h->yt[l], M * cp3::kMlpHidden)) return e;
            }
            for (int k = 0; k < 2; k++) {
                if (int e = dev_alloc(h, &h->dz[k], M * cp3::kMlpHidden)) return e;
                if (int e = dev_alloc(h, &h->dzt[k], M * cp3::kMlpHidden)) return e;
            }
            if (int e = dev_alloc(h, &h->loss_part, std::size_t(cp3::output_head_blocks(max_batch)))) return e;
        } else {
            for (int l = 0; l < 2; l++)
                if (int e = dev_alloc(h, &h->y[l], M * cp3::kMlpActLd)) return e;
        }
        if (int e = dev_alloc(h, &h->loss_dev, 1)) return e;
        CP3H_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->loss_host), sizeof(float)));
        CP3H_HIP(hipStreamSynchronize(h->stream));
        return CP3H_OK;
    }();
    if (rc != CP3H_OK) {
        cp3h_mlp_destroy(h);
        return rc;
    }
    *out = h;
    return CP3H_OK;
}

int cp3h_mlp_destroy(cp3h_mlp* h) {
    if (!h) return CP3H_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void* p : h->allocations) (void)hipFree(p);
    if (h->loss_host) (void)hipHostFree(h->loss_host);
    if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
    delete h;
    return CP3H_OK;
}

int cp3h_mlp_set_stream(cp3h_mlp* h, void* stream) {
    if (int rc = check(h)) return rc;
    CP3H_HIP(hipSetDevice(h->device));
    CP3H_HIP(hipStreamSynchronize(h->stream));
    h->stream = stream ? static_cast<hipStream_t>(stream) : h->own_stream;
    return CP3H_OK;
}

int cp3h_mlp_set_scaling(cp3h_mlp* h, const float* shift9, const float* scale9) {
    if (int rc = check(h)) return rc;
    if (!shift9 || !scale9) return fail(CP3H_ERR_ARGUMENT, "null scaling arrays");
    std::memcpy(h->scaling.shift, shift9, sizeof(float) * cp3::kMlpIn);
    std::memcpy(h->scaling.scale, scale9, sizeof(float) * cp3::kMlpIn);
    return CP3H_OK;
}

int cp3h_mlp_set_params(cp3h_mlp* h, const float* params, size_t n) {
    if (int rc = check(h)) return rc;
    if (!params || n != param_count()) return fail(CP3H_ERR_ARGUMENT, "parameter count mismatch");
    CP3H_HIP(hipSetDevice(h->device));
    CP3H_HIP(hipStreamSynchronize(h->stream));
    std::size_t off = 0;
    for (int l = 0; l < cp3::kMlpLayers; l++) {
        const LayerShape& s = kShape[l];
        std::vector<float> wpad(std::size_t(s.N) * s.K, 0.0f), bpad(s.N, 0.0f);
        for (int r = 0; r < s.n_valid; r++)
            for (int c = 0; c < s.k_valid; c++) wpad[std::size_t(r) * s.K + c] = params[off + std::size_t(r) * s.k_valid + c];
        off += std::size_t(s.n_valid) * s.k_valid;
        for (int r = 0; r < s.n_valid; r++) bpad[r] = params[off + r];
        off += s.n_valid;
        CP3H_HIP(hipMemcpy(h->w[l], wpad.data(), sizeof(float) * wpad.size(), hipMemcpyHostToDevice));
        CP3H_HIP(hipMemcpy(h->b[l], bpad.data(), sizeof(float) * bpad.size(), hipMemcpyHostToDevice));
        if (h->training) {
            const std::size_t nk = std::size_t(s.N) * s.K;
            CP3H_HIP(hipMemset(h->m1w[l], 0, sizeof(float) * nk));
            CP3H_HIP(hipMemset(h->m2w[l], 0, sizeof(float) * nk));
            CP3H_HIP(hipMemset(h->m1b[l], 0, sizeof(float) * s.N));
            CP3H_HIP(hipMemset(h->m2b[l], 0, sizeof(float) * s.N));
        }
    }
    h->adam_t = 0;
    if (int rc = refresh_bf16(h)) return rc;
    CP3H_HIP(hipStreamSynchronize(h->stream));
    return CP3H_OK;
}

int cp3h_mlp_get_params(cp3h_mlp* h, float* params, size_t n) {
    if (int rc = check(h)) return rc;
    if (!params || n != param_count()) return fail(CP3H_ERR_ARGUMENT, "parameter count mismatch");
    CP3H_HIP(hipSetDevice(h->device));
    CP3H_HIP(hipStreamSynchronize(h->stream));
    std::size_t off = 0;
    for (int l = 0; l < cp3::kMlpLayers; l++) {
        const LayerShape& s = kShape[l];
        std::vector<float> wpad(std::size_t(s.N) * s.K), bpad(s.N);
        CP3H_HIP(hipMemcpy(wpad.data(), h->w[l], sizeof(float) * wpad.size(), hipMemcpyDeviceToHost));
        CP3H_HIP(hipMemcpy(bpad.data(), h->b[l], sizeof(float) * bpad.size(), hipMemcpyDeviceToHost));
        for (int r = 0; r < s.n_valid; r++)
            for (int c = 0; c < s.k_valid; c++) params[off + std::size_t(r) * s.k_valid + c] = wpad[std::size_t(r) * s.K + c];
        off += std::size_t(s.n_valid) * s.k_valid;
        for (int r = 0; r < s.n_valid; r++) params[off + r] = bpad[r];
        off += s.n_valid;
    }
    return CP3H_OK;
}

int cp3h_mlp_forward(cp3h_mlp* h, const float* x_device, long long n, float* y_device) {
    if (int rc = check(h)) return rc;
    if (n < 0 || (n > 0 && (!x_device || !y_device))) return fail(CP3H_ERR_ARGUMENT, "bad forward arguments");
    CP3H_HIP(hipSetDevice(h->device));
    if (h->forward_kernel != CP3H_MLP_LAYERED) {
        CP3H_HIP(cp3::launch_mlp_forward_fused(x_device, n, y_device, h->wb, h->b, h->scaling, h->forward_kernel,
                                               nullptr, h->cu_count, 0, h->stream));
        return CP3H_OK;
    }
    for (long long r0 = 0; r0 < n; r0 += h->max_batch) {
        const int M = int(std::min<long long>(h->max_batch, n - r0));
        if (int rc = forward(h, x_device + r0 * cp3::kMlpIn, M, nullptr, y_device + r0 * cp3::kMlpOut)) return rc;
    }
    return CP3H_OK;
}

int cp3h_mlp_forward_grid(cp3h_mlp* h, const float* cos_device, int n_cos, const float* log10e_device, int n_bins,
                          const float* hyp_device, int n_hyp, const int* skip_below_device, float* y_device) {
    if (int rc = check(h)) return rc;
    if (n_cos < 0 || n_bins < 0 || n_hyp < 0) return fail(CP3H_ERR_ARGUMENT, "negative grid size");
    const long long n = (long long)n_cos * n_bins * n_hyp;
    if (n == 0) return CP3H_OK;
    if (!cos_device || !log10e_device || !hyp_device || !y_device)
        return fail(CP3H_ERR_ARGUMENT, "null grid arrays");
    CP3H_HIP(hipSetDevice(h->device));
    const cp3::MlpGrid g{cos_device, log10e_device, hyp_device, skip_below_device, nullptr, n_cos, n_bins};
    const int variant = h->forward_kernel == CP3H_MLP_LAYERED ? CP3H_MLP_FUSED4 : h->forward_kernel;
    CP3H_HIP(cp3::launch_mlp_forward_fused(nullptr, n, y_device, h->wb, h->b, h->scaling, variant, &g, h->cu_count, 0,
                                           h->stream));
    return CP3H_OK;
}

int cp3h_mlp_forward_grid_tiles(cp3h_mlp* h, const float* cos_device, int n_cos, const float* log10e_device,
                                int n_bins, const float* hyp_device, int n_hyp, const int* skip_below_device,
                                const long long* tiles_device, long long n_tiles, float* y_device) {
    if (int rc = check(h)) return rc;
    if (n_cos <= 0 || n_bins <= 0 || n_hyp < 0 || n_tiles < 0) return fail(CP3H_ERR_ARGUMENT, "bad grid size");
    if (n_tiles == 0 || n_hyp == 0) return CP3H_OK;
    if (!cos_device || !log10e_device || !hyp_device || !y_device || !tiles_device)
        return fail(CP3H_ERR_ARGUMENT, "null grid arrays");
    CP3H_HIP(hipSetDevice(h->device));
    const long long n = (long long)n_cos * n_bins * n_hyp;
    const cp3::MlpGrid g{cos_device, log10e_device, hyp_device, skip_below_device, tiles_device, n_cos, n_bins};
    const int variant = h->forward_kernel == CP3H_MLP_FUSED8 ? CP3H_MLP_FUSED8 : CP3H_MLP_FUSED4;
    CP3H_HIP(cp3::launch_mlp_forward_fused(nullptr, n, y_device, h->wb, h->b, h->scaling, variant, &g, h->cu_count,
                                           n_tiles, h->stream));
    return CP3H_OK;
}

int cp3h_mlp_set_forward_kernel(cp3h_mlp* h, int kernel) {
    if (int rc = check(h)) return rc;
    if (kernel != CP3H_MLP_LAYERED && kernel != CP3H_MLP_FUSED4 && kernel != CP3H_MLP_FUSED8 &&
        kernel != CP3H_MLP_FUSED64)
        return fail(CP3H_ERR_ARGUMENT, "unknown forward kernel");
    h->forward_kernel = kernel;
    return CP3H_OK;
}

int cp3h_mlp_gradients(cp3h_mlp* h, const float* x_device, const float* t_device, int M, float* grads, size_t n,
                       float* loss) {
    if (int rc = check(h)) return rc;
    if (int rc = check_batch(h, M)) return rc;
    if (!x_device || !t_device || !grads || n != param_count()) return fail(CP3H_ERR_ARGUMENT, "bad gradient arguments");
    CP3H_HIP(hipSetDevice(h->device));
    if (int rc = forward(h, x_device, M, t_device, nullptr)) return rc;
    if (int rc = backward(h, M)) return rc;
    if (int rc = loss_to_host(h, M, loss)) return rc;
    std::size_t off = 0;
    for (int l = 0; l < cp3::kMlpLayers; l++) {
        const LayerShape& s = kShape[l];
        std::vector<float> gpad(std::size_t(s.N) * s.K), bpad(s.N);
        CP3H_HIP(hipMemcpy(gpad.data(), h->gw[l], sizeof(float) * gpad.size(), hipMemcpyDeviceToHost));
        CP3H_HIP(hipMemcpy(bpad.data(), h->gb[l], sizeof(float) * bpad.size(), hipMemcpyDeviceToHost));
        for (int r = 0; r < s.n_valid; r++)
            for (int c = 0; c < s.k_valid; c++) grads[off + std::size_t(r) * s.k_valid + c] = gpad[std::size_t(r) * s.K + c];
        off += std::size_t(s.n_valid) * s.k_valid;
        for (int r = 0; r < s.n_valid; r++) grads[off + r] = bpad[r];
        off += s.n_valid;
    }
    return CP3H_OK;
}

int cp3h_mlp_train_step(cp3h_mlp* h, const float* x_device, const float* t_device, int M, float lr, float* loss) {
    if (int rc = check(h)) return rc;
    if (int rc = check_batch(h, M)) return rc;
    if (!x_device || !t_device) return fail(CP3H_ERR_ARGUMENT, "null batch");
    CP3H_HIP(hipSetDevice(h->device));
    if (int rc = forward(h, x_device, M, t_device, nullptr)) return rc;
    if (int rc = backward(h, M)) return rc;
    h->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const cp3::AdamStep a{lr, b1, b2, 1e-8f, float(1.0 / (1.0 - std::pow(double(b1), h->adam_t))),
                          float(1.0 / (1.0 - std::pow(double(b2), h->adam_t)))};
    for (int l = 0; l < cp3::kMlpLayers; l++) {
        const LayerShape& s = kShape[l];
        CP3H_HIP(cp3::launch_adam_layer(h->w[l], h->m1w[l], h->m2w[l], h->gw[l], s.N, s.K, s.n_valid, s.k_valid, a,
                                        h->wb[l], wb_ld(l), wb_dup(l), h->wtb[l], h->stream));
        CP3H_HIP(cp3::launch_adam_vec(h->b[l], h->m1b[l], h->m2b[l], h->gb[l], s.N, s.n_valid, a, h->stream));
    }
    if (loss) return loss_to_host(h, M, loss);
    return CP3H_OK;
}

int cp3h_mlp_sync(cp3h_mlp* h) {
    if (int rc = check(h)) return rc;
    CP3H_HIP(hipSetDevice(h->device));
    CP3H_HIP(hipStreamSynchronize(h->stream));
    return CP3H_OK;
}

}  // extern "C"
