// split16.hpp -- split-fp16 operand helpers of the matrix-core form-factor kernels.
//
// An FP32 value x is carried as x_h = fp16(x) and x_l = fp16(x - x_h) (both rounded to nearest;
// x - x_h is exact in FP32). A real product x y is taken as x_h y_h + x_l y_h + x_h y_l on
// v_mfma_f32_32x32x16_f16: the fp16 products are exact in the FP32 accumulators and x_l y_l is below
// FP32 rounding. Complex products pair (re, im) halves so that one B operand gives the real part of
// a complex product and its swapped, negated twin the imaginary part.
#pragma once

#include <hip/hip_runtime.h>

namespace hgx {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// accumulator register rr of lane (h = lane / 32) of a 32 x 32 MFMA tile holds tile row
// 8 (rr / 4) + 4 h + rr % 4
__device__ inline int acc_row(int rr, int h) { return 8 * (rr >> 2) + 4 * h + (rr & 3); }

// two FP32 values to fp16, rounded to nearest (v_cvt_pk_f16_f32)
__device__ inline f16x2 pk16(float a, float b) { return __builtin_convertvector((f32x2){a, b}, f16x2); }

// (a, -b) to fp16, rounded to nearest: one v_cvt_pk_f16_f32 with the negation as a source modifier
// (the compiler emits a separate sign flip for pk16(a, -b)); fp16(-b) = -fp16(b), so the pair equals
// pk16(a, -b) bit for bit. A register-only instruction: nothing is loaded or stored.
__device__ inline f16x2 pk16n(float a, float b) {
  f16x2 r;
  asm("v_cvt_pk_f16_f32 %0, %1, -%2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// The low halves of a pair whose high halves are h = pk16(a, b): (fp16(a - h.x), fp16(b - h.y)), each
// by one mixed-precision FMA (v_fma_mixlo_f16 / v_fma_mixhi_f16: an fp16 and two FP32 sources, an
// exact FP32 difference, one rounding to fp16), so the high halves are not converted back to FP32.
// Bit for bit the split of pk16(a - (float)h.x, b - (float)h.y). Register-only instructions.
__device__ inline f16x2 lo16(f16x2 h, float a, float b) {
  f16x2 r;
  asm("v_fma_mixlo_f16 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(a));
  asm("v_fma_mixhi_f16 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "+v"(r) : "v"(h), "v"(b));
  return r;
}

// The low halves of the swapped, negated pair (b, -a) from h = pk16(a, b): (fp16(b - h.y),
// fp16(h.x - a)), equal to swap_neg(lo16(h, a, b)) but for the sign of a zero half.
__device__ inline f16x2 lo16sn(f16x2 h, float a, float b) {
  f16x2 r;
  asm("v_fma_mixlo_f16 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(b));
  asm("v_fma_mixhi_f16 %0, %1, 1.0, -%2 op_sel_hi:[1,0,0]" : "+v"(r) : "v"(h), "v"(a));
  return r;
}

// (v.y, -v.x): the pair for the real part from the pair for the imaginary part (exact)
__device__ inline f16x2 swap_neg(f16x2 v) { return (f16x2){v[1], -v[0]}; }

__device__ inline f16x8 cat4(f16x2 a, f16x2 b, f16x2 c, f16x2 d) {
  typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
  f16x4 ab = __builtin_shufflevector(a, b, 0, 1, 2, 3);
  f16x4 cd = __builtin_shufflevector(c, d, 0, 1, 2, 3);
  return __builtin_shufflevector(ab, cd, 0, 1, 2, 3, 4, 5, 6, 7);
}

__device__ inline f32x16 mfma16(f16x8 a, f16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

// smallest e >= 0 with v <= 2^e (v = 0 gives 0)
__device__ inline int bound_exp(float v) {
  if (!(v > 0.0f)) return 0;
  int e;
  (void)frexpf(v, &e);  // v = m 2^e, m in [1/2, 1)
  return max(e, 0);
}

}  // namespace hgx
