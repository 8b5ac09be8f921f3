// Sine and cosine of a float for the emitted model and coupling functions.
//
// The argument is reduced by pi/2 (Cody-Waite): k = rint(x 2/pi) and
// r = x - k (C1 + C2 + C3) with pi/2 split into three floats, each product
// taken by a fused multiply-add, so |r| <= pi/4 (to within rounding).  Then
// sin r = r + r z (S1 + z (S2 + z S3)) and cos r = 1 - z/2 + z^2 (K1 + z (K2
// + z K3)), z = r^2, whose coefficients are least-squares fits on
// Chebyshev-distributed points of [0, pi/4] (relative error for sin,
// absolute for cos), rounded to float; the quadrant k mod 4 picks +-sin r or
// +-cos r.  Emulated in float32 with the same operations, the results are
// within 1.6 ulp of the correctly rounded sine and cosine for |x| <= 100.
// Larger and non-finite arguments take the library's sinf / cosf, whose
// full-range reduction (Payne-Hanek, 64-bit integer products) is what this
// avoids for the bounded arguments of the models (phase differences of
// wrapped oscillators), where it cost most of a step.
#ifndef TVBX_MATH_H
#define TVBX_MATH_H

__device__ __forceinline__ float tvbx_sincos_reduced(float x, int quarter) {
  const float kf = rintf(x * 0x1.45f306p-1f);  // 2 / pi
  float r = fmaf(-kf, 0x1.921fb6p+0f, x);      // C1
  r = fmaf(-kf, -0x1.777a5cp-25f, r);          // C2
  r = fmaf(-kf, -0x1.0p-49f, r);               // C3
  const float z = r * r;
  const float ps = fmaf(fmaf(-0x1.995346p-13f, z, 0x1.110772p-7f), z,
                        -0x1.555546p-3f);
  const float sn = fmaf(r * z, ps, r);
  const float pc = fmaf(fmaf(0x1.9a0168p-16f, z, -0x1.6c0c88p-10f), z,
                        0x1.55554ap-5f);
  const float cs = fmaf(z * z, pc, fmaf(-0.5f, z, 1.0f));
  const int q = ((int)kf + quarter) & 3;
  const float v = (q & 1) ? cs : sn;
  return (q & 2) ? -v : v;
}

// the library's functions out of line: called only for the rare large or
// non-finite argument, they stay out of the unrolled loops' code
__device__ __noinline__ float tvbx_sinf_full(float x) { return sinf(x); }
__device__ __noinline__ float tvbx_cosf_full(float x) { return cosf(x); }

__device__ __forceinline__ float tvbx_sinf(float x) {
  if (!(fabsf(x) <= 100.0f)) return tvbx_sinf_full(x);
  return tvbx_sincos_reduced(x, 0);
}

__device__ __forceinline__ float tvbx_cosf(float x) {
  if (!(fabsf(x) <= 100.0f)) return tvbx_cosf_full(x);
  return tvbx_sincos_reduced(x, 1);
}

// The same without the branch, for a group of coupling terms whose range
// is checked once (the step kernels' gather groups): NaN outside |x| <= 100,
// which makes the group's sum NaN and has it redone through tvbx_sinf /
// tvbx_cosf.  With no branch between them, the sines of a group interleave.
__device__ __forceinline__ float tvbx_sinf_nb(float x) {
  const float v = tvbx_sincos_reduced(x, 0);
  return fabsf(x) <= 100.0f ? v : __builtin_nanf("");
}

__device__ __forceinline__ float tvbx_cosf_nb(float x) {
  const float v = tvbx_sincos_reduced(x, 1);
  return fabsf(x) <= 100.0f ? v : __builtin_nanf("");
}

// a / b within one ulp (faithful rounding): the reciprocal's estimate
// (v_rcp_f32, about one ulp) times a, corrected once by the residual,
// q = a r, q + r (a - q b) -- four instructions against the ten or so of
// the correctly rounded division.  The first quotient is returned as it is
// when it or b is not finite (b = 0, b = +-inf -- where a / b is +-0 but the
// residual would be 0 * inf -- and NaN operands).
__device__ __forceinline__ float tvbx_divf(float a, float b) {
  const float r = __builtin_amdgcn_rcpf(b);
  const float q = a * r;
  const float e = fmaf(-q, b, a);
  return (__builtin_isfinite(q) && __builtin_isfinite(b)) ? fmaf(e, r, q)
                                                            : q;
}

#endif
