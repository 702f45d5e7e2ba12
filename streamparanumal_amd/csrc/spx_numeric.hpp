// Storage-type conversions for the reduced-precision variants.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace spx {

// bf16 is carried as its 16 raw bits.
struct bf16_t {
  uint16_t bits;
};

// NaN of any sign and payload -> 0x7fc0, as torch's bfloat16 conversion.
constexpr uint16_t kBf16NaN = 0x7fc0;

// f64 -> bf16, round to nearest even, without the double rounding of a
// f64->f32->bf16 chain: round to f32 with round-to-odd (exact when the f32
// keeps 16 spare bits beyond bf16), then round-to-nearest-even to bf16.  A
// value past the f32 range rounds to f32 infinity, whose magnitude exceeds
// it, so round-to-odd takes the largest finite f32 instead, and that rounds
// to bf16 infinity (it lies above the midpoint between the largest bf16 and
// 2^128), as the exact value does.
__device__ __forceinline__ uint16_t f64_to_bf16_bits(double x) {
  const float f = (float)x;
  uint32_t u = __float_as_uint(f);
  const double back = (double)f;
  u -= fabs(back) > fabs(x) ? 1u : 0u;  // truncate toward zero (sign-magnitude)
  u |= back != x ? 1u : 0u;             // sticky bit: round to odd
  // nearest even to bf16: one v_cvt_pk_bf16_f32 on gfx950
  const uint16_t r = __builtin_bit_cast(uint16_t, (__bf16)__uint_as_float(u));
  return x != x ? kBf16NaN : r;
}

// f32 -> bf16, round to nearest even
__device__ __forceinline__ uint16_t f32_to_bf16_bits(float f) {
  return f != f ? kBf16NaN : __builtin_bit_cast(uint16_t, (__bf16)f);
}

__device__ __forceinline__ float bf16_bits_to_f32(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

// Storage traits: load as float / double, store from float / double.
template <typename S> struct Store;
template <> struct Store<double> {
  static __device__ __forceinline__ double to_f64(double v) { return v; }
  static __device__ __forceinline__ float to_f32(double v) { return (float)v; }
  static __device__ __forceinline__ double from_f64(double v) { return v; }
  static __device__ __forceinline__ double from_f32(float v) { return (double)v; }
};
template <> struct Store<float> {
  static __device__ __forceinline__ double to_f64(float v) { return (double)v; }
  static __device__ __forceinline__ float to_f32(float v) { return v; }
  static __device__ __forceinline__ float from_f64(double v) { return (float)v; }
  static __device__ __forceinline__ float from_f32(float v) { return v; }
};
template <> struct Store<bf16_t> {
  static __device__ __forceinline__ double to_f64(bf16_t v) { return (double)bf16_bits_to_f32(v.bits); }
  static __device__ __forceinline__ float to_f32(bf16_t v) { return bf16_bits_to_f32(v.bits); }
  static __device__ __forceinline__ bf16_t from_f64(double v) { return bf16_t{f64_to_bf16_bits(v)}; }
  static __device__ __forceinline__ bf16_t from_f32(float v) { return bf16_t{f32_to_bf16_bits(v)}; }
};

}  // namespace spx
