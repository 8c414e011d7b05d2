// FP16C: 16-bit storage of x = (f - w_q) / w_q with one sign bit, a 4-bit
// exponent (bias 14) and an 11-bit mantissa: one more mantissa bit than
// IEEE binary16 for |x| < 4 (values beyond saturate), subnormal spacing 2^-24
// below 2^-13.  Low-Mach populations stay well inside |x| < 1.
//
// The codec works through one scaling: the float whose bits are the code's
// magnitude shifted left by 12 (exponent field = the code's exponent,
// mantissa = its 11 mantissa bits followed by zeros) is exactly 2^-113 times
// the code's value.  For codes below 2^-13 that float is an FP32 denormal,
// whose spacing 2^-149 is 2^-36 in x, 2^-12 of the subnormal code step, so
// the same statement covers both ranges.  Decode is a shift, a mask and one
// multiply by 2^113 (folded into the caller's multiply by w); stochastic
// encode is one multiply by 2^-113, an integer add of 12 random bits below
// the kept ones and a shift.  Both rely on FP32 denormals being kept, which
// is the default for device code (the kernel descriptors say so) and for
// host code without flush-to-zero.
//
// All functions here are host and device code, so the CPU tests can run the
// device codec itself on every code and on dense sets of values.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fl {
namespace lbm {

struct fp16c {
    uint16_t bits;
};

constexpr float kFp16cUp = 0x1p113f;     // code float -> x
constexpr float kFp16cDown = 0x1p-113f;  // x -> code float

// 2^-113 times the value of code c (sign kept): sign-extending the 16-bit
// code and shifting left by 12 copies the sign into bits 27..31; the mask
// keeps bit 31 and the magnitude in bits 12..26.
__host__ __device__ inline float fp16c_code_float(uint16_t c) {
    const uint32_t s = (uint32_t)(int32_t)(int16_t)c << 12;
    return __builtin_bit_cast(float, s & 0x87FFF000u);
}

__host__ __device__ inline float fp16c_decode(uint16_t c) { return fp16c_code_float(c) * kFp16cUp; }

// Stochastic rounding of x to FP16C given xs = x * 2^-113 (callers fold the
// scale into a multiply they already do): the magnitude bits of xs are the
// code times 2^12 plus the dropped fraction, so adding 12 random bits (the
// low 12 of rnd) and shifting rounds up with probability equal to that
// fraction.  Below 2^-13 the product xs itself rounds x to the nearest
// multiple of 2^-36 first.  Inf, NaN and |x| >= 4 saturate to the largest
// code (the sum stays below 2^32).
__host__ __device__ inline uint16_t fp16c_encode_sr_scaled(float xs, uint32_t rnd) {
    const uint32_t sum = (__builtin_bit_cast(uint32_t, fabsf(xs)) + (rnd & 0xFFFu)) >> 12;
    const uint32_t code = sum < 0x7FFFu ? sum : 0x7FFFu;
    return (uint16_t)(((__builtin_bit_cast(uint32_t, xs) >> 16) & 0x8000u) | code);
}

// Round to nearest, ties to even, exact in both ranges (set-up and fills;
// the collide stores through the stochastic form).
__host__ __device__ inline uint16_t fp16c_encode_rne(float x) {
    const uint32_t b = __builtin_bit_cast(uint32_t, x);
    const uint32_t sign = (b >> 16) & 0x8000u;
    const uint32_t mag = b & 0x7FFFFFFFu;
    uint32_t code;
    if (mag >= 0x39000000u) {  // |x| >= 2^-13
        const uint32_t m2 = mag + 0x7FFu + ((mag >> 12) & 1u);
        code = (mag >= 0x7F800000u || m2 >= 0x40800000u) ? 0x7FFFu : (m2 - (113u << 23)) >> 12;
    } else {
        code = (uint32_t)rintf(__builtin_bit_cast(float, mag) * 16777216.0f);  // 2^24, exact scaling
    }
    return (uint16_t)(sign | code);
}

}  // namespace lbm
}  // namespace fl
