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
// evaluated as x / (1 + 2^a) with a = -u log2(e) = x (kGeluA0 + kGeluA1 x^2): one hardware
// exp2 and one reciprocal.  For x -> -inf, 2^a overflows to +inf and the reciprocal gives 0;
// for x -> +inf it gives 1.
constexpr float kGeluU = 1.5957691216057308f;               // 2 sqrt(2/pi)
constexpr float kGeluC = 0.044715f;
constexpr float kGeluA0 = -2.3022082010588f;                // -2 sqrt(2/pi) log2(e)
constexpr float kGeluA1 = -2.3022082010588f * 0.044715f;

__device__ inline float gelu_sigmoid(float x) {
    const float a = x * __builtin_fmaf(kGeluA1, x * x, kGeluA0);
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(a));
}

__device__ inline float gelu(float x) { return x * gelu_sigmoid(x); }

// d gelu / dx = s + x s (1 - s) du/dx, s = sigmoid(u), du/dx = 2 sqrt(2/pi) (1 + 3 * 0.044715 x^2)
__device__ inline float gelu_grad(float x) {
    const float s = gelu_sigmoid(x);
    const float du = kGeluU * __builtin_fmaf(3.0f * kGeluC * x, x, 1.0f);
    return __builtin_fmaf(x * s * (1.0f - s), du, s);
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Two activations as two-term bf16: hi = bf16(a), lo = bf16(a - hi), each pair packed in 32 bits
// (first value in the low half), with packed conversions.
__device__ inline void two_term_pair(float a0, float a1, unsigned& hi, unsigned& lo) {
    hi = __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2{a0, a1}), bf16x2));
    const float h0 = __builtin_bit_cast(float, hi << 16), h1 = __builtin_bit_cast(float, hi & 0xffff0000u);
    lo = __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2{a0 - h0, a1 - h1}), bf16x2));
}

}  // namespace cp3
