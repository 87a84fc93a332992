// Device helpers shared by the surrogate network's training (mlp.hip) and fused inference
// (mlp_infer.hip) kernels: operand vector types and the activation function.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

namespace cp3 {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// GELU in its tanh form (PyTorch's gelu(approximate="tanh")):
//   gelu(x) = 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))) = x * sigmoid(u),
//   u = 2 sqrt(2/pi) (x + 0.044715 x^3),
// evaluated with one hardware exp and one reciprocal.  For u -> -inf the exp overflows to
// +inf and the reciprocal gives 0; for u -> +inf it gives 1.
constexpr float kGeluU = 1.5957691216057308f;   // 2 sqrt(2/pi)
constexpr float kGeluC = 0.044715f;

__device__ inline float gelu_arg(float x) { return kGeluU * x * __builtin_fmaf(kGeluC * x, x, 1.0f); }

__device__ inline float gelu_sigmoid(float x) { return __builtin_amdgcn_rcpf(1.0f + __expf(-gelu_arg(x))); }

__device__ inline float gelu(float x) { return x * gelu_sigmoid(x); }

// d gelu / dx = s + x s (1 - s) du/dx, s = sigmoid(u), du/dx = 2 sqrt(2/pi) (1 + 3 * 0.044715 x^2)
__device__ inline float gelu_grad(float x) {
    const float s = gelu_sigmoid(x);
    const float du = kGeluU * __builtin_fmaf(3.0f * kGeluC * x, x, 1.0f);
    return __builtin_fmaf(x * s * (1.0f - s), du, s);
}

}  // namespace cp3
