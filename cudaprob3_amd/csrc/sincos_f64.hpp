// sin and cos of one FP64 argument for the FP64 pair kernels: the per-pair phases
// e^{i phi} = cos phi + i sin phi of the reference's getA (physics.hpp:462-474), which it takes
// from the C library.
//
// Cody and Waite's reduction: k = rint(x * 2/pi) quarter turns, r = x - k pi/2 with pi/2 split
// into three doubles (each the double nearest what the previous parts leave), every product
// subtracted by one fused multiply-add, so r carries the error of its last rounding only for
// |k| far beyond the phases a layer reaches; then the Taylor polynomials of sin and cos on
// |r| <= pi/4 to degrees 17 and 16 (truncation below 1e-19 and 2e-18 there), in Horner form,
// and the quadrant k mod 4 applied by swapping and negating.  Within 1.5 ulp of the exact values
// and 1 ulp of the C library's (the CPU tests measure both), with no large-argument path:
// |x| > kSincosMaxArg is left to the caller's library call.  In the FP64 pair kernel it took
// 4-5 % off the kernel time against the library sincos.
//
// Host and device: the CPU tests compile this header with g++ against the C library.
#pragma once

#if defined(__HIPCC__)
#define CP3_HOST_DEVICE __host__ __device__
#else
#define CP3_HOST_DEVICE
#endif

#include <cstdint>

namespace cp3 {

constexpr double kHalfPi1 = 0x1.921fb54442d18p+0;      // pi/2 = kHalfPi1 + kHalfPi2 + kHalfPi3
constexpr double kHalfPi2 = 0x1.1a62633145c07p-54;
constexpr double kHalfPi3 = -0x1.f1976b7ed8fbcp-110;
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;    // 2/pi
constexpr double kSincosMaxArg = 1.0e12;                // |k| < 2^40: the reduction stays exact enough

// sin r and cos r for |r| <= pi/4 (a little beyond is fine): Taylor coefficients (-1)^n / (2n+1)!
// and (-1)^n / (2n)!.
CP3_HOST_DEVICE inline void sincos_quarter(double r, double& s, double& c) {
    const double r2 = r * r;
    double ps = 2.8114572543455206e-15;                       //  1/17!
    ps = __builtin_fma(r2, ps, -7.647163731819816e-13);       // -1/15!
    ps = __builtin_fma(r2, ps, 1.6059043836821613e-10);       //  1/13!
    ps = __builtin_fma(r2, ps, -2.505210838544172e-08);       // -1/11!
    ps = __builtin_fma(r2, ps, 2.7557319223985893e-06);       //  1/9!
    ps = __builtin_fma(r2, ps, -1.984126984126984e-04);       // -1/7!
    ps = __builtin_fma(r2, ps, 8.333333333333333e-03);        //  1/5!
    ps = __builtin_fma(r2, ps, -1.6666666666666666e-01);      // -1/3!
    s = __builtin_fma(r * r2, ps, r);
    double pc = 4.779477332387385e-14;                        //  1/16!
    pc = __builtin_fma(r2, pc, -1.1470745597729725e-11);      // -1/14!
    pc = __builtin_fma(r2, pc, 2.08767569878681e-09);         //  1/12!
    pc = __builtin_fma(r2, pc, -2.755731922398589e-07);       // -1/10!
    pc = __builtin_fma(r2, pc, 2.48015873015873e-05);         //  1/8!
    pc = __builtin_fma(r2, pc, -1.388888888888889e-03);       // -1/6!
    pc = __builtin_fma(r2, pc, 4.1666666666666664e-02);       //  1/4!
    pc = __builtin_fma(r2, pc, -0.5);                         // -1/2!
    c = __builtin_fma(r2, pc, 1.0);
}

// sin x and cos x; returns false (s, c untouched) when |x| > kSincosMaxArg or x is not finite.
CP3_HOST_DEVICE inline bool sincos_reduced_f64(double x, double& s, double& c) {
    if (!(__builtin_fabs(x) <= kSincosMaxArg)) return false;
    const double k = __builtin_rint(x * kTwoOverPi);
    double r = __builtin_fma(-k, kHalfPi1, x);
    r = __builtin_fma(-k, kHalfPi2, r);
    r = __builtin_fma(-k, kHalfPi3, r);
    double sr, cr;
    sincos_quarter(r, sr, cr);
    // k mod 4 from the last two mantissa bits of k + 1.5 * 2^52 (an integer of [2^52, 2^53), where
    // those bits are k's own modulo 4 for |k| < 2^51, negative k included)
    const unsigned q = unsigned(__builtin_bit_cast(std::uint64_t, k + 6755399441055744.0)) & 3u;
    const double sa = (q & 1u) ? cr : sr;
    const double ca = (q & 1u) ? sr : cr;
    s = (q & 2u) ? -sa : sa;
    c = ((q + 1u) & 2u) ? -ca : ca;
    return true;
}

}  // namespace cp3
