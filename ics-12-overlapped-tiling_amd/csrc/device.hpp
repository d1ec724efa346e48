// device.hpp — per-point update formulas for every stencil (device code).
//
// Each formula keeps the association order of the reference's CPU loop
// (or, where the reference has no CPU loop, of the first-time-step equations of
// its generated kernel text), so that with FP contraction disabled the GPU
// result equals a CPU evaluation of the same expression bit for bit.
// OpenCL's rsqrt (approximate in the reference's kernels) is replaced by the
// correctly rounded 1.0f/sqrtf(x), which is what the reference's CPU loop
// computes (gradient-2d.cpp:697-699).
#pragma once
#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

namespace ot {
namespace dev {

// 1.0f / sqrtf(x) as the IEEE expression (both operations correctly rounded),
// the reference value of every rsqrt below.
__device__ __forceinline__ float rsq_ieee(float x) { return 1.0f / __builtin_sqrtf(x); }

// The same value from the hardware reciprocal-square-root and reciprocal
// estimates, for the arguments the stencils produce (x >= 1e-20: every rsqrt
// argument is a sum with a positive epsilon of 1e-4, 1e-8, 1e-9 or 1e-20):
//   s = RN(sqrt(x)): from the v_rsq_f32 estimate y0 ~ 1/sqrt(x), the product
//     s0 = x y0 corrected once by s = s0 + (y0/2)(x - s0^2), the residual
//     exact through FMA (P. Markstein, "Computation of elementary functions on
//     the IBM RISC System/6000 processor", IBM J. Res. Dev. 34(1), 1990);
//   RN(1/s): the v_rcp_f32 estimate with one correction y' = y + y(1 - s y),
//     again with the residual exact through FMA;
//   x = +inf gives 0, NaN stays NaN.
// Neither correction is correctly rounded for every input in theory; both
// are on this hardware for every float of [1e-20, +inf], which
// ot_selftest_rsq checks bit for bit against the IEEE expression. 8 VALU
// instructions and 2 transcendental ones, against 26 for the IEEE expression
// (which also guards denormal and zero inputs). Starting the reciprocal from
// y0 instead of v_rcp_f32 misses 196 inputs (s just below a power of two,
// where the correction from below ends on a tie and rounds to even).
__device__ __forceinline__ float rsq(float x) {
  const float y0 = __builtin_amdgcn_rsqf(x);
  const float s0 = x * y0;
  const float s = __builtin_fmaf(__builtin_fmaf(-s0, s0, x), 0.5f * y0, s0);
  float y = __builtin_amdgcn_rcpf(s);
  y = __builtin_fmaf(__builtin_fmaf(-s, y, 1.0f), y, y);
  return (x == __builtin_inff()) ? 0.0f : y;
}
__device__ __forceinline__ float sq(float x) { return x * x; }

// x * m under the DX9 rule (v_mul_legacy_f32, one VALU instruction like
// v_mul_f32; clang exposes no builtin for it): a zero operand of either sign
// gives +0 whatever the other one is (infinite and NaN included), and two
// nonzero operands give the IEEE product, denormals kept. It lets a Dirichlet
// mask ride on an update's final multiply (m = 0 outside the domain) and still
// give exactly the +0 that the reference keeps on its ring, also when the
// masked sum is negative, infinite or NaN. Inside (m != 0) it differs from the
// IEEE product only for x = -0, which it turns into +0.
__device__ __forceinline__ float mul_legacy(float x, float m) {
  float r;
  asm("v_mul_legacy_f32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(m));
  return r;
}

// Wavefront x-neighbours for the wave-per-tile kernels (lanes = consecutive x
// columns): DPP wave_shr:1 / wave_shl:1 with bound_ctrl, folded by the compiler
// into the consuming VALU instruction.
// lane i receives lane i-1's value (lane 0 receives 0)
__device__ __forceinline__ float from_left(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, false));
}
// lane i receives lane i+1's value (lane 63 receives 0)
__device__ __forceinline__ float from_right(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xF, 0xF, false));
}

// jacobi-1d.cpp:633-635
__device__ __forceinline__ float j1d3(float l, float c, float r) { return 0.333f * ((l + c) + r); }
// jacobi-1d-5pt.cpp:507-511
__device__ __forceinline__ float j1d5(float m2, float m1, float c, float p1, float p2) {
  return 0.2f * ((((m2 + m1) + c) + p1) + p2);
}
// jacobi-1d-7pt.cpp:515-521 (0.2f kept as in the reference)
__device__ __forceinline__ float j1d7(float m3, float m2, float m1, float c, float p1, float p2,
                                      float p3) {
  return 0.2f * ((((((m3 + m2) + m1) + c) + p1) + p2) + p3);
}
// jacobi-2d.cpp:824-828 (1.0f/5.0f == 0.2f): left, centre, right, top(y-1), bottom(y+1)
__device__ __forceinline__ float j2d(float l, float c, float r, float t, float b) {
  return 0.2f * ((((l + c) + r) + t) + b);
}
// poisson-2d.cpp:644-648
__device__ __forceinline__ float p2d(float tl, float tc, float tr, float ml, float mc, float mr,
                                     float bl, float bc, float br) {
  return 8.0f * mc - (((((((tl + tc) + tr) + ml) + mr) + bl) + bc) + br);
}
// gradient-2d.cpp:691-699 (float path)
__device__ __forceinline__ float g2d(float l, float c, float r, float t, float b) {
  return c + rsq((((0.0001f + sq(c - l)) + sq(c - r)) + sq(c - t)) + sq(c - b));
}
// jacobi-3d.cpp:734-741: order (y-1, c, y+1, x-1, x+1, z+1, z-1); j3d() is
// 0.143f times this sum
__device__ __forceinline__ float j3d_sum(float xm, float xp, float ym, float yp, float zm, float zp,
                                         float c) {
  return (((((ym + c) + yp) + xm) + xp) + zp) + zm;
}
__device__ __forceinline__ float j3d(float xm, float xp, float ym, float yp, float zm, float zp,
                                     float c) {
  return 0.143f * j3d_sum(xm, xp, ym, yp, zm, zp, c);
}
// gradient-3d.cpp:305-307: val0..val6 = x-1, c, x+1, y-1, y+1, z-1, z+1
__device__ __forceinline__ float g3d(float xm, float xp, float ym, float yp, float zm, float zp,
                                     float c) {
  return c + rsq((((((0.0001f + sq(c - xm)) + sq(c - xp)) + sq(c - ym)) + sq(c - yp)) + sq(c - zm)) +
                 sq(c - zp));
}
// Poisson-Jacobi relaxation with right-hand side f and unit spacing:
// u' = (sum6 - f) / 6 (no reference driver; the reference's poisson-3d has no RHS)
__device__ __forceinline__ float p3d7(float xm, float xp, float ym, float yp, float zm, float zp,
                                      float f) {
  return ((((((xm + xp) + ym) + yp) + zm) + zp) - f) / 6.0f;
}
// cs-3d.cpp:286-308 (kernel constants :161-163; eps taken as the float 1e-9f,
// pow(d, 2.0f) as d*d). Arguments by offset (dz, dy, dx) from the centre u0;
// each tap is the address of the same-named value in the on-chip steps
// (cs-3d.cpp:325-379), where val_0_m1_0 is row y-1 and val_0_p1_0 row y+1 (the
// first step of the reference's time tile reads those two the other way round,
// cs-3d.cpp:258-263; not reproduced). c1..c4 weight the centre's forward
// neighbours and its z-1, y-1, x-1 neighbours by the centre value over the
// forward-difference gradient magnitude at the centre and at those neighbours.
__device__ __forceinline__ float cs3d(float u0, float zm, float zp, float ym, float yp, float xm, float xp,
                                      float zm_yp, float zm_xp, float zp_ym, float ym_xp, float zp_xm,
                                      float yp_xm) {
  const float eps = 1e-9f, sum_aref = 1.1f, alpha = 1.1f;
  const float c1 = (u0 * rsq(((eps + sq(zp - u0)) + sq(yp - u0)) + sq(xp - u0))) / sum_aref;
  const float c2 = (u0 * rsq(((eps + sq(u0 - zm)) + sq(zm_yp - zm)) + sq(zm_xp - zm))) / sum_aref;
  const float c3 = (u0 * rsq(((eps + sq(zp_ym - ym)) + sq(u0 - ym)) + sq(ym_xp - ym))) / sum_aref;
  const float c4 = (u0 * rsq(((eps + sq(zp_xm - xm)) + sq(yp_xm - xm)) + sq(u0 - xm))) / sum_aref;
  return ((((alpha * u0 + c1 * ((zp + yp) + xp)) + c2 * zm) + c3 * ym) + c4 * xm) /
         ((((alpha + 3.0f * c1) + c2) + c3) + c4);
}

// The same update from the rsqrt of the four gradient sums, computed once per
// point by the caller: g0 at the centre, gz, gy, gx at its z-1, y-1, x-1
// neighbours (each the same expression and order as in cs3d above).
__device__ __forceinline__ float cs3d_g(float u0, float zm, float zp, float ym, float yp, float xm, float xp,
                                        float g0, float gz, float gy, float gx) {
  const float sum_aref = 1.1f, alpha = 1.1f;
  const float c1 = (u0 * g0) / sum_aref;
  const float c2 = (u0 * gz) / sum_aref;
  const float c3 = (u0 * gy) / sum_aref;
  const float c4 = (u0 * gx) / sum_aref;
  return ((((alpha * u0 + c1 * ((zp + yp) + xp)) + c2 * zm) + c3 * ym) + c4 * xm) /
         ((((alpha + 3.0f * c1) + c2) + c3) + c4);
}

// cdsc-tv-update-2d.cpp:290-309. c3 duplicates c1 as in the reference (its
// MATLAB comment repeats c1's formula); V and F read-only, alpha a parameter.
__device__ __forceinline__ float tv2d(float tc, float tr, float ml, float mc, float mr, float bl,
                                      float bc, float v, float f, float alpha) {
  const float c1 = (mc * rsq((1e-8f + sq(bc - mc)) + sq(mr - mc))) / v;
  const float c2 = (mc * rsq((1e-8f + sq(mc - tc)) + sq(tr - tc))) / v;
  const float c3 = (mc * rsq((1e-8f + sq(bc - mc)) + sq(mr - mc))) / v;
  const float c4 = (mc * rsq((1e-8f + sq(bl - ml)) + sq(mc - ml))) / v;
  return (1.0f / ((((alpha + c1) + c2) + c3) + c4)) *
         ((((alpha * f + c1 * bc) + c2 * tc) + c3 * mr) + c4 * ml);
}

// The same update from the rsqrt of the gradient sums, computed once per point
// by the caller: g0 at the centre (c1 and its duplicate c3), gt at the top
// neighbour (c2), gl at the left neighbour (c4), each the sum tv2d forms.
__device__ __forceinline__ float tv2d_g(float tc, float ml, float mc, float mr, float bc, float v, float f,
                                        float alpha, float g0, float gt, float gl) {
  const float c1 = (mc * g0) / v;
  const float c2 = (mc * gt) / v;
  const float c3 = c1;
  const float c4 = (mc * gl) / v;
  return (1.0f / ((((alpha + c1) + c2) + c3) + c4)) *
         ((((alpha * f + c1 * bc) + c2 * tc) + c3 * mr) + c4 * ml);
}

// rician-denoise-2d.cpp:154-162, :271-281, :322-327.
struct RicianK {
  float dt, sigma2, gamma;
};
__device__ __forceinline__ float rician_g(float c, float l, float r, float t, float b) {
  return rsq((((1.0e-20f + sq(c - r)) + sq(c - l)) + sq(c - b)) + sq(c - t));
}
__device__ __forceinline__ float rician_u(float c, float f, float ul, float ur, float ut, float ub,
                                          float gl, float gr, float gt, float gb, RicianK k) {
  float rr = (c * f) / k.sigma2;
  rr = (rr * (2.38944f + rr * (0.950037f + rr))) / (4.65314f + rr * (2.57541f + rr * (1.48937f + rr)));
  return (c + k.dt * ((((ub * gb + ut * gt) + ur * gr) + ul * gl) + (k.gamma * f) * rr)) /
         (1.0f + k.dt * ((((gr + gl) + gb) + gt) + k.gamma));
}

// fdtd-2d.cpp:281-283, :290-292, :320-322
__device__ __forceinline__ float fdtd_ey(float ey, float hz, float hz_top) { return ey - 0.5f * (hz - hz_top); }
__device__ __forceinline__ float fdtd_ex(float ex, float hz, float hz_left) { return ex - 0.5f * (hz - hz_left); }
__device__ __forceinline__ float fdtd_hz(float hz, float ex, float ex_right, float ey, float ey_bottom,
                                         bool physical) {
  return physical ? hz - 0.7f * (((ex_right - ex) + ey_bottom) - ey)
                  : hz - 0.7f * (((ex_right - ex) + ey_bottom) + ey);
}

// Map a linear workgroup id to a logical tile id so that the workgroups the
// hardware places on one XCD (ids congruent mod 8) get a contiguous run of
// tiles, keeping shared halos within that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
  const int xcd = b & 7, local = b >> 3;
  const int q = n >> 3, rem = n & 7;
  return xcd * q + (xcd < rem ? xcd : rem) + local;
}

}  // namespace dev
}  // namespace ot
