// stencil.hpp -- per-cell formulas of the level-set path, written so that the
// FP64 build reproduces the reference's rounding: same operations, same
// association order, no FMA contraction (the library is compiled with
// -ffp-contract=off in double precision).  Shared by the drop-in kernels, the
// narrow-band kernels and the FP64 fused stage kernels.
#pragma once

#include "lsmhip_core.hpp"

namespace lsmhip {
namespace stencil {

// One HJ-WENO5 reconstruction from five scaled differences, in the order of
// src/toolbox/lsm_spatial_derivatives3d.f.in:1296-1323.  tiny is the additive
// term of eps: the templated 1e-36, or the literal 1e-35 of the 2D narrow-band
// routine (lsm_spatial_derivatives2d_local.f.in:892).
template <class T>
__device__ __forceinline__ T weno5(T v1, T v2, T v3, T v4, T v5, T tiny = T(1.e-36)) {
  const T one_third = T(1.0 / 3.0), seven_sixths = T(7.0 / 6.0), eleven_sixths = T(11.0 / 6.0);
  const T one_sixth = T(1.0 / 6.0), five_sixths = T(5.0 / 6.0);
  const T thirteen_twelfths = T(13.0 / 12.0), one_fourth = T(0.25);
  T m = v1 * v1;
  m = v2 * v2 > m ? v2 * v2 : m;
  m = v3 * v3 > m ? v3 * v3 : m;
  m = v4 * v4 > m ? v4 * v4 : m;
  m = v5 * v5 > m ? v5 * v5 : m;
  const T eps = T(1e-6) * m + tiny;
  const T p1 = one_third * v1 - seven_sixths * v2 + eleven_sixths * v3;
  const T p2 = -(one_sixth * v2) + five_sixths * v3 + one_third * v4;
  const T p3 = one_third * v3 + five_sixths * v4 - one_sixth * v5;
  const T S1 = thirteen_twelfths * sq(v1 - T(2) * v2 + v3) + one_fourth * sq(v1 - T(4) * v2 + T(3) * v3);
  const T S2 = thirteen_twelfths * sq(v2 - T(2) * v3 + v4) + one_fourth * sq(v2 - v4);
  const T S3 = thirteen_twelfths * sq(v3 - T(2) * v4 + v5) + one_fourth * sq(T(3) * v3 - T(4) * v4 + v5);
  T a1 = T(0.1) / sq(S1 + eps);
  T a2 = T(0.6) / sq(S2 + eps);
  T a3 = T(0.3) / sq(S3 + eps);
  const T inv = T(1.0) / (a1 + a2 + a3);
  a1 = a1 * inv;
  a2 = a2 * inv;
  a3 = a3 * inv;
  return a1 * p1 + a2 * p2 + a3 * p3;
}

// In FP32 the rearranged reconstruction of the fused stages (fastf::weno5
// below: same weights and candidates, fewer instructions, different rounding)
// serves every WENO5 routine; the additive tiny term only matters when every
// |v| < 1e-15, where all candidates vanish.
namespace fastf {
__device__ float weno5(float v1, float v2, float v3, float v4, float v5);
}
__device__ __forceinline__ float weno5(float v1, float v2, float v3, float v4, float v5, float tiny = 1.e-36f) {
  (void)tiny;
  return fastf::weno5(v1, v2, v3, v4, v5);
}

// Undivided differences from seven samples f[s+3] = phi(offset s) along an axis:
// D1(s) = f(s) - f(s-1) (face s-1/2 stored at s), D2(s) = -(D1(s) - D1(s+1)),
// D3(s) = D2(s) - D2(s-1)  (lsmComputeDn, lsm_spatial_derivatives3d.f.in:49-175).
#define LSM_D1(s) (f[(s) + 3] - f[(s) + 2])

template <class T>
__device__ __forceinline__ T eno2_plus(const T* d1, const T* d2) {  // d1[s+2], d2[s+2]
  return fabs(d2[2]) < fabs(d2[3]) ? d1[3] - T(0.5) * d2[2] : d1[3] - T(0.5) * d2[3];
}
template <class T>
__device__ __forceinline__ T eno2_minus(const T* d1, const T* d2) {
  return fabs(d2[1]) < fabs(d2[2]) ? d1[2] + T(0.5) * d2[1] : d1[2] + T(0.5) * d2[2];
}
// ENO3, lsm_spatial_derivatives3d.f.in:837-903 (one rounding per accumulation).
template <class T>
__device__ __forceinline__ T eno3_plus(const T* d1, const T* d2, const T* d3) {
  const T third = T(1.0 / 3.0), sixth = T(1.0 / 6.0);
  T r = d1[3];
  if (fabs(d2[2]) < fabs(d2[3])) {
    r = r - T(0.5) * d2[2];
    r = fabs(d3[2]) < fabs(d3[3]) ? r - sixth * d3[2] : r - sixth * d3[3];
  } else {
    r = r - T(0.5) * d2[3];
    r = fabs(d3[3]) < fabs(d3[4]) ? r + third * d3[3] : r + third * d3[4];
  }
  return r;
}
template <class T>
__device__ __forceinline__ T eno3_minus(const T* d1, const T* d2, const T* d3) {
  const T third = T(1.0 / 3.0), sixth = T(1.0 / 6.0);
  T r = d1[2];
  if (fabs(d2[1]) < fabs(d2[2])) {
    r = r + T(0.5) * d2[1];
    r = fabs(d3[1]) < fabs(d3[2]) ? r + third * d3[1] : r + third * d3[2];
  } else {
    r = r + T(0.5) * d2[2];
    r = fabs(d3[2]) < fabs(d3[3]) ? r - sixth * d3[2] : r - sixth * d3[3];
  }
  return r;
}

// Differences indexed [s+2] for s = -2..3 (d1), -2..2 (d2), -1..2 (d3 at [s+2]).
template <class T>
__device__ __forceinline__ void differences(const T f[7], T d1[6], T d2[5], T d3[5]) {
#pragma unroll
  for (int s = -2; s <= 3; s++) d1[s + 2] = LSM_D1(s);
#pragma unroll
  for (int s = -2; s <= 2; s++) d2[s + 2] = -(d1[s + 2] - d1[s + 3]);
#pragma unroll
  for (int s = -1; s <= 2; s++) d3[s + 2] = d2[s + 2] - d2[s + 1];
}

// Plus and minus one-sided derivatives of order 1, 2, 3 (ENO) or 5 (WENO).
template <int ORDER, class T>
__device__ __forceinline__ void plus_minus(const T f[7], T inv, T& p, T& m) {
  if (ORDER == 1) {
    p = LSM_D1(1) * inv;
    m = LSM_D1(0) * inv;
  } else if (ORDER == 2) {
    T d1[6], d2[5];
#pragma unroll
    for (int s = -1; s <= 2; s++) d1[s + 2] = LSM_D1(s);
#pragma unroll
    for (int s = -1; s <= 1; s++) d2[s + 2] = -(d1[s + 2] - d1[s + 3]);
    p = eno2_plus(d1, d2) * inv;
    m = eno2_minus(d1, d2) * inv;
  } else if (ORDER == 3) {
    T d1[6], d2[5], d3[5];
    differences(f, d1, d2, d3);
    p = eno3_plus(d1, d2, d3) * inv;
    m = eno3_minus(d1, d2, d3) * inv;
  } else {
    T v[6];
#pragma unroll
    for (int s = -2; s <= 3; s++) v[s + 2] = LSM_D1(s) * inv;
    p = weno5(v[5], v[4], v[3], v[2], v[1]);
    m = weno5(v[0], v[1], v[2], v[3], v[4]);
  }
}

// Velocity-upwinded derivative: v > 0 takes the minus (backward) side.
template <int ORDER, class T>
__device__ __forceinline__ T upwind(const T f[7], T inv, bool vpos) {
  if (ORDER == 1) return (vpos ? LSM_D1(0) : LSM_D1(1)) * inv;
  if (ORDER == 2) {
    T d1[6], d2[5];
#pragma unroll
    for (int s = -1; s <= 2; s++) d1[s + 2] = LSM_D1(s);
#pragma unroll
    for (int s = -1; s <= 1; s++) d2[s + 2] = -(d1[s + 2] - d1[s + 3]);
    return (vpos ? eno2_minus(d1, d2) : eno2_plus(d1, d2)) * inv;
  }
  if (ORDER == 3) {
    T d1[6], d2[5], d3[5];
    differences(f, d1, d2, d3);
    return (vpos ? eno3_minus(d1, d2, d3) : eno3_plus(d1, d2, d3)) * inv;
  }
  if (vpos) return weno5(LSM_D1(-2) * inv, LSM_D1(-1) * inv, LSM_D1(0) * inv, LSM_D1(1) * inv, LSM_D1(2) * inv);
  return weno5(LSM_D1(3) * inv, LSM_D1(2) * inv, LSM_D1(1) * inv, LSM_D1(0) * inv, LSM_D1(-1) * inv);
}
#undef LSM_D1

// |grad phi| with per-component and result thresholds
// (lsm_spatial_derivatives3d.f.in:3622-3650 uses 1e-8; the local twin uses zero_tol).
template <class T>
__device__ __forceinline__ T grad_magnitude(T x, T y, T z, T tol) {
  T sx = x * x, sy = y * y, sz = z * z;
  if (sx < tol) sx = T(0);
  if (sy < tol) sy = T(0);
  if (sz < tol) sz = T(0);
  const T t = sqrt(sx + sy + sz);
  return t < tol ? T(0) : t;
}

// Godunov |grad phi|^2 for normal speed of sign vpos
// (lsm_level_set_evolution3d.f.in:210-224).
template <class T>
__device__ __forceinline__ T godunov_sq(bool vpos, T xp, T yp, T zp, T xm, T ym, T zm) {
  const T z0 = T(0);
  if (vpos)
    return maxr(sq(maxr(xm, z0)), sq(minr(xp, z0))) + maxr(sq(maxr(ym, z0)), sq(minr(yp, z0))) +
           maxr(sq(maxr(zm, z0)), sq(minr(zp, z0)));
  return maxr(sq(minr(xm, z0)), sq(maxr(xp, z0))) + maxr(sq(minr(ym, z0)), sq(maxr(yp, z0))) +
         maxr(sq(minr(zm, z0)), sq(maxr(zp, z0)));
}

// Reinitialization RHS of one cell, lsm_reinitialization3d.f.in:167-199 (and :251 for phi0).
template <class T>
__device__ __forceinline__ T reinit_rhs(T s, bool use_phi0, T xp, T yp, T zp, T xm, T ym, T zm, T dx_sq) {
  const T gp[3] = {xp, yp, zp}, gm[3] = {xm, ym, zm};
  T star[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    T p, m;
    if (s > T(0)) {
      p = maxr(-gp[d], T(0));
      m = maxr(gm[d], T(0));
    } else {
      p = maxr(gp[d], T(0));
      m = maxr(-gm[d], T(0));
    }
    star[d] = maxr(p, m);
  }
  if (fabs(s) >= kZeroTol) {
    const T n2 = star[0] * star[0] + star[1] * star[1] + star[2] * star[2];
    const T sgn = use_phi0 ? s / sqrt(s * s + dx_sq) : s / sqrt(s * s + n2 * dx_sq);
    return sgn * (T(1) - sqrt(n2));
  }
  return T(0);
}

// Curvature numerator shared by lsm_level_set_evolution3d.f.in:433-451 and
// lsm_curvature3d.f.in:103-120.
template <class T>
__device__ __forceinline__ T curv_num(T xx, T xy, T xz, T yy, T yz, T zz, T x, T y, T z) {
  return xx * y * y + yy * x * x - T(2) * xy * x * y + xx * z * z + zz * x * x - T(2) * xz * x * z + yy * z * z +
         zz * y * y - T(2) * yz * y * z;
}

// Gaussian-curvature numerator, lsm_curvature3d.f.in:230-237 (and :490-497
// with the x-z mixed derivative taken from phi_z along x); the narrow-band
// routines of lsm_curvature3d_local.f.in:258-264, 408-414, 555-561 use the same
// expression.
template <class T>
__device__ __forceinline__ T gauss_num(T x, T y, T z, T xx, T yy, T zz, T xy, T xz, T yz) {
  return x * x * (yy * zz - yz * yz) + y * y * (xx * zz - xz * xz) + z * z * (xx * yy - xy * xy) +
         T(2) * (x * y * (xz * yz - xy * zz) + y * z * (xy * xz - yz * xx) + x * z * (xy * yz - xz * yy));
}

// x^1.5 for x >= 1e-11.  FP64: the exact product x * sqrt(x) carried as a
// double-double (sqrt's remainder x - s^2 and the product's rounding error are
// both exact with FMA) and rounded once, i.e. correctly rounded except within
// about 2^-100 of a rounding boundary -- the value a correctly rounded pow(x,
// 1.5) returns, which is what the C library's pow gives the reference's
// compiled Fortran.  FP32: x * sqrt(x).
template <class T>
__device__ __forceinline__ T pow_three_halves(T x) {
  if constexpr (sizeof(T) == 8) {
    const double s = sqrt(x);
    const double p = x * s;
    if (__builtin_isinf(p)) return p;
    const double e = fma(x, s, -p);       // x s - p
    const double r = fma(-s, s, x);       // x - s^2
    const double t = r / (2.0 * s);       // sqrt(x) - s to second order
    return p + fma(x, t, e);
  } else {
    return x * sqrt(x);
  }
}

// Per-axis Godunov choice for external + normal motion,
// lsm_level_set_evolution3d.f.in:629-693.
template <class T>
__device__ __forceinline__ T ext_normal_pick(T u, T vn, T pp, T pm) {
  const T hp = u + vn * pp, hm = u + vn * pm;
  if (hp > T(0)) return hm > T(0) ? pm : -(u / vn);
  if (hm < T(0)) return pp;
  return fabs(hp) > fabs(hm) ? pp : pm;
}

// Normal-motion CFL quantity |V| |grad|_1-weighted / |grad|, lsm_math_utils3d.f.in:413-429.
template <class T>
__device__ __forceinline__ T normal_h_over_dx(T absv, T xp, T yp, T zp, T xm, T ym, T zm, T mdx2, T ix, T iy, T iz,
                                              T* norm_out = nullptr) {
  const T x = maxr(fabs(xp), fabs(xm)), y = maxr(fabs(yp), fabs(ym)), z = maxr(fabs(zp), fabs(zm));
  const T n = sqrt(x * x + y * y + z * z + mdx2);
  if (norm_out) *norm_out = n;
  return absv / n * (x * ix + y * iy + z * iz);
}

// ---------------------------------------------------------------------------
// FP32 forms: the same HJ-WENO5 scheme rearranged for fewer VALU instructions
// (FP32 results are held to an error bound against FP64, not to bitwise
// equality).
//  * The reconstruction is evaluated on undivided differences and scaled by
//    1/dx once: the weights are invariant under a common scaling of v.
//  * With m = max(v_k^2) the weights alpha_k = d_k / (S_k + 1e-6 m)^2 equal
//    d_k / (S_k/m + 1e-6)^2 up to a common factor; S_k/m lies in [0, 34], so
//    the normalised terms stay far from FP32 overflow and underflow.  (The
//    reference's additive 1e-36 only matters when every |v| < 1e-15, where all
//    candidate stencils vanish.)
//  * The three weights are put over a common denominator, so the reconstruction
//    needs two hardware reciprocals (v_rcp_f32, 1 ulp) instead of four IEEE
//    divisions.
// ---------------------------------------------------------------------------
namespace fastf {

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// One WENO5 stencil (v1..v5) in two parts.  win5 holds what depends on the
// five values as a set -- the normalised smoothness products and the
// candidate differences; fwd finishes weno5(v1, .., v5) and rev finishes
// weno5(v5, .., v1), whose smoothness indicators are those of the forward
// order taken in reverse.  The plus derivative at a cell and the minus
// derivative at the next cell along an axis use the same five differences in
// opposite orders, so one window serves both.
//
// The nonlinear weights are unchanged when the smoothness indicators S_k and
// eps = 1e-6 max v^2 are scaled together: win5 divides them by max v^2
// ((max |v|)^2, exact as rounding is monotonic) and multiplies by 4, so that
// 4 S_k = (13/3) a_k^2 + b_k^2 with a_k, b_k the reference's two terms.  With
// alpha_k = d_k / e_k^2 over the common denominator e1^2 e2^2 e3^2 and
// p1 - p2 = (a1 - a2)/3, p3 - p2 = (a2 - a3)/6,
//   sum_k w_k p_k = p2 + (0.1 q23 (a1 - a2)/3 + 0.3 q12 (a2 - a3)/6) / sum_k d_k q_(other two).
struct Win {
  float q12, q13, q23, d12, d23;
};

__device__ __forceinline__ Win win5(float v1, float v2, float v3, float v4, float v5) {
  const float ma = fmaxf(fmaxf(fmaxf(fmaxf(fabsf(v1), fabsf(v2)), fabsf(v3)), fabsf(v4)), fabsf(v5));
  const float im = rcp(fmaf(ma, ma, 1e-30f));
  const float a1 = fmaf(-2.f, v2, v1) + v3, b1 = fmaf(3.f, v3, fmaf(-4.f, v2, v1));
  const float a2 = fmaf(-2.f, v3, v2) + v4, b2 = v2 - v4;
  const float a3 = fmaf(-2.f, v4, v3) + v5, b3 = fmaf(3.f, v3, fmaf(-4.f, v4, v5));
  const float e1 = fmaf(fmaf(a1 * (13.f / 3.f), a1, b1 * b1), im, 4e-6f);
  const float e2 = fmaf(fmaf(a2 * (13.f / 3.f), a2, b2 * b2), im, 4e-6f);
  const float e3 = fmaf(fmaf(a3 * (13.f / 3.f), a3, b3 * b3), im, 4e-6f);
  const float q1 = e1 * e1, q2 = e2 * e2, q3 = e3 * e3;
  return Win{q1 * q2, q1 * q3, q2 * q3, a1 - a2, a2 - a3};
}
// weno5(v1, .., v5) from win5(v1, .., v5)
__device__ __forceinline__ float fwd(const Win& w, float v2, float v3, float v4) {
  const float den = fmaf(0.1f, w.q23, fmaf(0.6f, w.q13, 0.3f * w.q12));
  const float num = fmaf(w.q23, w.d12 * (0.1f / 3.f), w.q12 * (w.d23 * 0.05f));
  const float p2 = fmaf(-1.f / 6.f, v2, fmaf(5.f / 6.f, v3, (1.f / 3.f) * v4));
  return fmaf(num, rcp(den), p2);
}
// weno5(v5, .., v1) from win5(v1, .., v5): S1 and S3 trade places, and the
// candidate differences become -(a2 - a3) and -(a1 - a2)
__device__ __forceinline__ float rev(const Win& w, float v2, float v3, float v4) {
  const float den = fmaf(0.1f, w.q12, fmaf(0.6f, w.q13, 0.3f * w.q23));
  const float num = -fmaf(w.q12, w.d23 * (0.1f / 3.f), w.q23 * (w.d12 * 0.05f));
  const float p2 = fmaf(-1.f / 6.f, v4, fmaf(5.f / 6.f, v3, (1.f / 3.f) * v2));
  return fmaf(num, rcp(den), p2);
}

__device__ __forceinline__ float weno5(float v1, float v2, float v3, float v4, float v5) {
  return fwd(win5(v1, v2, v3, v4, v5), v2, v3, v4);
}
// fwd and rev of one window together (the minus derivative at a cell and the
// plus derivative at the cell before it): the two denominators share 0.6 q13,
// the numerators the products q23 d12 and q12 d23, the linear parts (5/6) v3.
// The same weights and candidates as fwd and rev, rounded in another order.
__device__ __forceinline__ void fwd_rev(const Win& w, float v2, float v3, float v4, float& f, float& r) {
  const float m13 = 0.6f * w.q13;
  const float den_f = fmaf(0.1f, w.q23, fmaf(0.3f, w.q12, m13));
  const float den_r = fmaf(0.1f, w.q12, fmaf(0.3f, w.q23, m13));
  const float pa = w.q23 * w.d12, pb = w.q12 * w.d23;
  const float num_f = fmaf(0.1f / 3.f, pa, 0.05f * pb);
  const float num_r = -fmaf(0.1f / 3.f, pb, 0.05f * pa);
  const float t3 = (5.f / 6.f) * v3;
  const float p2_f = fmaf(-1.f / 6.f, v2, fmaf(1.f / 3.f, v4, t3));
  const float p2_r = fmaf(-1.f / 6.f, v4, fmaf(1.f / 3.f, v2, t3));
  f = fmaf(num_f, rcp(den_f), p2_f);
  r = fmaf(num_r, rcp(den_r), p2_r);
}

// Plus and minus WENO5 derivatives from seven samples f[s+3] along an axis:
// minus uses D(-2..2), plus uses D(3..-1) = the window D(-1..3) reversed.
__device__ __forceinline__ void plus_minus(const float f[7], float inv, float& p, float& m) {
  float d[6];
#pragma unroll
  for (int s = -2; s <= 3; s++) d[s + 2] = f[s + 3] - f[s + 2];
  p = rev(win5(d[1], d[2], d[3], d[4], d[5]), d[2], d[3], d[4]) * inv;
  m = fwd(win5(d[0], d[1], d[2], d[3], d[4]), d[1], d[2], d[3]) * inv;
}

// Velocity-upwinded WENO5 derivative (vpos: backward/minus stencil).
__device__ __forceinline__ float upwind(const float f[7], float inv, bool vpos) {
  float d[6];
#pragma unroll
  for (int s = -2; s <= 3; s++) d[s + 2] = f[s + 3] - f[s + 2];
  const float a = vpos ? d[0] : d[5], b = vpos ? d[1] : d[4], c = vpos ? d[2] : d[3];
  const float e = vpos ? d[3] : d[2], g = vpos ? d[4] : d[1];
  return weno5(a, b, c, e, g) * inv;
}

// upwind() where the side taken is usually the same across a wave (a smooth
// velocity): a wave whose lanes all take one side forms that side's window
// alone, without a per-lane choice of operands.  Each lane's arithmetic is
// that of upwind(), so a cell's result does not depend on the other lanes.
__device__ __forceinline__ float upwind_wave(const float f[7], float inv, bool vpos) {
  const unsigned long long pos = __builtin_amdgcn_ballot_w64(vpos), live = __builtin_amdgcn_ballot_w64(true);
  float d[6];
#pragma unroll
  for (int s = -2; s <= 3; s++) d[s + 2] = f[s + 3] - f[s + 2];
  // (the empty asm keeps each side's arithmetic whole in its own branch; merged
  // tails would need register moves on every path)
  float r;
  if (pos == live) {
    r = fwd(win5(d[0], d[1], d[2], d[3], d[4]), d[1], d[2], d[3]) * inv;
    asm volatile("" : "+v"(r));
  } else if (pos == 0) {
    r = fwd(win5(d[5], d[4], d[3], d[2], d[1]), d[4], d[3], d[2]) * inv;
    asm volatile("" : "+v"(r));
  } else {
    r = upwind(f, inv, vpos);
  }
  return r;
}

// Reinitialization RHS of one cell (the reference's formula,
// lsm_reinitialization3d.f.in:167-199 and :251) with the hardware square root
// and reciprocal square root (about 1 ulp each) in place of the correctly
// rounded FP32 sqrt and division sequences.
__device__ __forceinline__ float reinit_rhs(float s, bool use_phi0, float xp, float yp, float zp, float xm, float ym,
                                            float zm, float dx_sq) {
  const float gp[3] = {xp, yp, zp}, gm[3] = {xm, ym, zm};
  float n2 = 0.f;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const float p = s > 0.f ? fmaxf(-gp[d], 0.f) : fmaxf(gp[d], 0.f);
    const float m = s > 0.f ? fmaxf(gm[d], 0.f) : fmaxf(-gm[d], 0.f);
    const float st = fmaxf(p, m);
    n2 = fmaf(st, st, n2);
  }
  if (!(fabsf(s) >= kZeroTol)) return 0.f;
  const float sgn = s * __builtin_amdgcn_rsqf(fmaf(s, s, use_phi0 ? dx_sq : n2 * dx_sq));
  return sgn * (1.f - __builtin_amdgcn_sqrtf(n2));
}

}  // namespace fastf

}  // namespace stencil
}  // namespace lsmhip
