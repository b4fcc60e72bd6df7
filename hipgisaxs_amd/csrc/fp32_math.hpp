// fp32_math.hpp -- FP32 device math for the FP32C paths on gfx950.
//
// Angles are carried in turns (1 turn = 2 pi rad) because v_sin_f32 / v_cos_f32 take turns and
// reduce the argument themselves (measured max abs error 1.25e-7 for |t| <= 250 turns). v_sin_f32
// also keeps its relative accuracy at small arguments (measured max relative error 8.7e-7 from 1e-12
// to 1 turn, a near-constant 4.9e-7 below 1e-4 turns), so a quotient of two sines (the Dirichlet
// kernel next to Bragg peaks) takes it directly.
#pragma once

#include <hip/hip_runtime.h>

namespace hgx {

constexpr float kTwoPiF = 6.28318530717958647692f;
constexpr float kPiF = 3.14159265358979323846f;

__device__ inline float sin_turn(float t) { return __builtin_amdgcn_sinf(t); }
__device__ inline float cos_turn(float t) { return __builtin_amdgcn_cosf(t); }

// J0 and J1 of a real argument, Abramowitz & Stegun, Handbook of Mathematical Functions,
// eqs. 9.4.1, 9.4.3 (J0) and 9.4.4, 9.4.6 (J1): |error| < 1e-7 absolute (checked against scipy).
// j1_over_x returns J1(x)/x without the division for |x| <= 3.
__device__ inline float j1_over_x(float x) {
  float ax = fabsf(x);
  if (ax <= 3.0f) {
    float y = (x * (1.0f / 3.0f)) * (x * (1.0f / 3.0f));
    return 0.5f + y * (-0.56249985f + y * (0.21093573f + y * (-0.03954289f + y * (0.00443319f +
           y * (-0.00031761f + y * 0.00001109f)))));
  }
  float z = 3.0f / ax;
  float f1 = 0.79788456f + z * (0.00000156f + z * (0.01659667f + z * (0.00017105f + z * (-0.00249511f +
             z * (0.00113653f + z * (-0.00020033f))))));
  float t1 = ax - 2.35619449f + z * (0.12499612f + z * (0.00005650f + z * (-0.00637879f + z * (0.00074348f +
             z * (0.00079824f + z * (-0.00029166f))))));
  // cos of t1 (radians) through the turn-based unit
  return f1 * cos_turn(t1 * (1.0f / kTwoPiF)) * rsqrtf(ax) / ax;
}

__device__ inline float j0f_as(float x) {
  float ax = fabsf(x);
  if (ax <= 3.0f) {
    float y = (x * (1.0f / 3.0f)) * (x * (1.0f / 3.0f));
    return 1.0f + y * (-2.2499997f + y * (1.2656208f + y * (-0.3163866f + y * (0.0444479f +
           y * (-0.0039444f + y * 0.0002100f)))));
  }
  float z = 3.0f / ax;
  float f0 = 0.79788456f + z * (-0.00000077f + z * (-0.00552740f + z * (-0.00009512f + z * (0.00137237f +
             z * (-0.00072805f + z * 0.00014476f)))));
  float t0 = ax - 0.78539816f + z * (-0.04166397f + z * (-0.00003954f + z * (0.00262573f + z * (-0.00054125f +
             z * (-0.00029333f + z * 0.00013558f)))));
  return f0 * cos_turn(t0 * (1.0f / kTwoPiF)) * rsqrtf(ax);
}

struct cf {
  float r, i;
};
__device__ inline cf operator+(cf a, cf b) { return cf{a.r + b.r, a.i + b.i}; }
__device__ inline cf operator*(cf a, cf b) { return cf{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ inline cf operator*(float s, cf a) { return cf{s * a.r, s * a.i}; }

// exp(i 2 pi t) for t in turns
__device__ inline cf expi_turn(float t) { return cf{cos_turn(t), sin_turn(t)}; }

// Dirichlet kernel D_N(x) = sin(N x / 2) / sin(x / 2) for x already wrapped to (-pi, pi] and
// given in turns (u = x / 2 pi in (-1/2, 1/2]); the x -> 0 limit is N. Both sines are the hardware's
// (relatively accurate at small arguments, see above; near a peak their common bias cancels in the
// quotient). The quotient takes the hardware reciprocal (v_rcp_f32, 1 ulp) instead of the IEEE
// division sequence: den is a sine, never near the float range limits.
__device__ inline float dirichlet_turn(float u, float N) {
  const float den = sin_turn(0.5f * u);
  const float v = sin_turn(0.5f * (N * u)) * __builtin_amdgcn_rcpf(den);
  return fabsf(den) < 1e-30f ? N : v;
}

}  // namespace hgx
