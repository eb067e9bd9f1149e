// adjoint3d.hip -- reverse mode of one TVD-RK stage of normal motion,
//   out = a*u_c + b*(u_s + dt*L(u_s)),  L_i = -V_i sqrt(G_i),
// where G_i is the Godunov |grad phi|^2 of the six HJ-WENO5 one-sided
// derivatives of u_s at cell i (the forward of lsmhip_weno5_stage_normal and
// of the chain HJ_WENO5 -> ADD_[CONST_]NORMAL_VEL_TERM -> TVD_RK3_STAGEn,
// src/toolbox/lsm_spatial_derivatives3d.f.in:1161-1611,
// src/toolbox/lsm_level_set_evolution3d.f.in:131-349,
// src/toolbox/lsm_tvd_runge_kutta3d.f:269-500).  No counterpart in the
// reference, which has no derivatives of its time stepping.
//
// Given lam_out = dJ/d out on the fill box, the stage adds
//   lam_s += b*lam_out + (b dt dL/du_s)^T lam_out   (over the ghost box),
//   lam_c += a*lam_out,
//   lam_V += -b dt sqrt(G) lam_out                  (when V is a field).
// The transpose of dL/du_s is formed in two passes.  Pass 1 (one thread per
// fill-box cell i) differentiates L_i with respect to its 19-point stencil:
// per axis the plus and minus WENO5 values depend on u_s(i-3 .. i+3) through
// six undivided differences, and the Godunov selection picks which of them
// enters G_i; the seven resulting weights per axis, times b dt lam_out(i), go
// to a caller-provided scratch of 21 fields.  Pass 2 (one thread per ghost-box
// cell j) gathers the 21 weights that name j, so no atomics are needed and the
// sums are deterministic.
//
// Derivatives follow the implemented functions: the WENO5 weights
// a_k = d_k/(S_k + eps)^2 with eps = 1e-6 max v^2 (the max contributes through
// its argmax), the Godunov max/min selections (the branch the forward's
// maxr takes, the plus branch on ties), and sqrt(G) (zero derivative where
// G = 0).
#include "lsmhip_core.hpp"
#include "reduce.hpp"
#include "stencil.hpp"

using namespace lsmhip;

namespace {

constexpr int kQ = 21;  // 3 axes x 7 stencil offsets

// Value of one WENO5 reconstruction and its gradient with respect to v1..v5.
// W is homogeneous of degree 1 in v (eps = 1e-6 max v^2 scales with v^2), so
// its gradient is homogeneous of degree 0 and is evaluated on v / max|v|: the
// weights then stay in range in FP32 whatever the scale of v (the FP32 forward
// normalises the same way).  The 1e-36 floor of eps is below FP64 rounding of
// the normalised eps and is left out; v = 0 gives the linear-weight limit.
template <class T>
__device__ __forceinline__ T weno5_grad(const T vin[5], T g[5]) {
  int am = 0;
  T ma = fabs(vin[0]);
#pragma unroll
  for (int k = 1; k < 5; k++)
    if (fabs(vin[k]) > ma) {
      ma = fabs(vin[k]);
      am = k;
    }
  if (ma == T(0)) {  // d/dv of the linear-weight combination 0.1 p1 + 0.6 p2 + 0.3 p3
    g[0] = T(0.1 / 3.0);
    g[1] = T(-0.7 / 6.0 - 0.1);
    g[2] = T(1.1 / 6.0 + 0.5 + 0.1);
    g[3] = T(0.2 + 0.25);
    g[4] = T(-0.05);
    return T(0);
  }
  const T sc = T(1) / ma;
  T v[5];
#pragma unroll
  for (int k = 0; k < 5; k++) v[k] = vin[k] * sc;
  const T eps = T(1e-6) * (v[am] * v[am]);
  const T a1 = v[0] - T(2) * v[1] + v[2], b1 = v[0] - T(4) * v[1] + T(3) * v[2];
  const T a2 = v[1] - T(2) * v[2] + v[3], b2 = v[1] - v[3];
  const T a3 = v[2] - T(2) * v[3] + v[4], b3 = T(3) * v[2] - T(4) * v[3] + v[4];
  const T c13 = T(13.0 / 12.0);
  const T S1 = c13 * a1 * a1 + T(0.25) * b1 * b1;
  const T S2 = c13 * a2 * a2 + T(0.25) * b2 * b2;
  const T S3 = c13 * a3 * a3 + T(0.25) * b3 * b3;
  const T p1 = T(1.0 / 3.0) * v[0] - T(7.0 / 6.0) * v[1] + T(11.0 / 6.0) * v[2];
  const T p2 = -T(1.0 / 6.0) * v[1] + T(5.0 / 6.0) * v[2] + T(1.0 / 3.0) * v[3];
  const T p3 = T(1.0 / 3.0) * v[2] + T(5.0 / 6.0) * v[3] - T(1.0 / 6.0) * v[4];
  const T e1 = S1 + eps, e2 = S2 + eps, e3 = S3 + eps;
  const T al1 = T(0.1) / (e1 * e1), al2 = T(0.6) / (e2 * e2), al3 = T(0.3) / (e3 * e3);
  const T inv = T(1) / (al1 + al2 + al3);
  const T w1 = al1 * inv, w2 = al2 * inv, w3 = al3 * inv;
  const T W = w1 * p1 + w2 * p2 + w3 * p3;
  // dW/dS_k = -2 w_k (p_k - W) / (S_k + eps); eps enters every S_k
  const T t1 = T(-2) * w1 * (p1 - W) / e1, t2 = T(-2) * w2 * (p2 - W) / e2, t3 = T(-2) * w3 * (p3 - W) / e3;
  const T s6 = T(13.0 / 6.0), s3 = T(13.0 / 3.0);
  g[0] = w1 * T(1.0 / 3.0) + t1 * (s6 * a1 + T(0.5) * b1);
  g[1] = w1 * T(-7.0 / 6.0) + w2 * T(-1.0 / 6.0) + t1 * (-s3 * a1 - T(2) * b1) + t2 * (s6 * a2 + T(0.5) * b2);
  g[2] = w1 * T(11.0 / 6.0) + w2 * T(5.0 / 6.0) + w3 * T(1.0 / 3.0) + t1 * (s6 * a1 + T(1.5) * b1) +
         t2 * (-s3 * a2) + t3 * (s6 * a3 + T(1.5) * b3);
  g[3] = w2 * T(1.0 / 3.0) + w3 * T(5.0 / 6.0) + t2 * (s6 * a2 - T(0.5) * b2) + t3 * (-s3 * a3 - T(2) * b3);
  g[4] = w3 * T(-1.0 / 6.0) + t3 * (s6 * a3 + T(0.5) * b3);
  g[am] += (t1 + t2 + t3) * T(2e-6) * v[am];
  return W * ma;
}

struct AdjArgs {
  const real_t* us;
  const real_t* lam_out;
  real_t* lam_us;
  real_t* lam_uc;
  real_t* lam_v;
  const real_t* vnf;
  real_t vn;
  real_t a, b, dt;
  real_t inv[3];
  real_t* q;  // kQ fields over the ghost box
  Box gb, fb;
  long long vol;
  int sy, sz;
};

__device__ __forceinline__ long long goff(const AdjArgs& A, int i, int j, int k) {
  return (long long)(i - A.gb.ilo) + (long long)A.sy * (j - A.gb.jlo) + (long long)A.sz * (k - A.gb.klo);
}

__global__ void __launch_bounds__(256) adj_weights_kernel(AdjArgs A) {
  LSMHIP_CELL(A.fb);
  const long long o = goff(A, i, j, k);
  const real_t lam = A.lam_out[o];
  const real_t v = A.vnf ? A.vnf[o] : A.vn;
  const real_t coef = A.b * A.dt * lam;
  if (A.lam_uc && A.a != real_t(0)) A.lam_uc[o] += A.a * lam;
  const bool moving = fabs(v) >= kZeroTol;
  if (coef == real_t(0) || !moving) {
#pragma unroll
    for (int q = 0; q < kQ; q++) A.q[q * A.vol + o] = real_t(0);
    return;
  }
  const long long step[3] = {1, A.sy, A.sz};
  real_t gp[3][5], gm[3][5], pp[3], pm[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    real_t f[7];
#pragma unroll
    for (int s = -3; s <= 3; s++) f[s + 3] = A.us[o + s * step[d]];
    real_t D[6];  // D[t+2] = f(t) - f(t-1), t = -2..3
#pragma unroll
    for (int t = -2; t <= 3; t++) D[t + 2] = f[t + 3] - f[t + 2];
    const real_t h = A.inv[d];
    real_t vp[5] = {D[5] * h, D[4] * h, D[3] * h, D[2] * h, D[1] * h};
    real_t vm[5] = {D[0] * h, D[1] * h, D[2] * h, D[3] * h, D[4] * h};
    pp[d] = weno5_grad(vp, gp[d]);
    pm[d] = weno5_grad(vm, gm[d]);
  }
  // Godunov selection per axis: the selected one-sided value enters G squared
  const bool vpos = v > real_t(0);
  real_t G = real_t(0), selp[3], selm[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const real_t xm = vpos ? maxr(pm[d], real_t(0)) : minr(pm[d], real_t(0));
    const real_t xp = vpos ? minr(pp[d], real_t(0)) : maxr(pp[d], real_t(0));
    if (xm * xm > xp * xp) {  // maxr of the forward: the plus branch on ties
      G += xm * xm;
      selm[d] = xm;
      selp[d] = real_t(0);
    } else {
      G += xp * xp;
      selm[d] = real_t(0);
      selp[d] = xp;
    }
  }
  const real_t sg = sqrt(G);
  if (A.lam_v && A.vnf) A.lam_v[o] += -A.b * A.dt * sg * lam;
  // dL/dp = -V sel / sqrt(G), bounded by |V| since |sel| <= sqrt(G); zero where G vanishes
  const real_t rs = sg > real_t(0) ? real_t(1) / sg : real_t(0);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const real_t cp = coef * (-v) * (selp[d] * rs) * A.inv[d], cm = coef * (-v) * (selm[d] * rs) * A.inv[d];
    real_t w[7] = {0, 0, 0, 0, 0, 0, 0};  // w[s+3], s = -3..3
#pragma unroll
    for (int kk = 0; kk < 5; kk++) {
      const int tp = 3 - kk;  // plus: v_k = D(4-k), k = kk+1
      w[tp + 3] += cp * gp[d][kk];
      w[tp + 2] -= cp * gp[d][kk];
      const int tm = kk - 2;  // minus: v_k = D(k-3)
      w[tm + 3] += cm * gm[d][kk];
      w[tm + 2] -= cm * gm[d][kk];
    }
#pragma unroll
    for (int s = 0; s < 7; s++) A.q[(d * 7 + s) * A.vol + o] = w[s];
  }
}

__global__ void __launch_bounds__(256) adj_gather_kernel(AdjArgs A) {
  LSMHIP_CELL(A.gb);
  const long long o = goff(A, i, j, k);
  const Box& f = A.fb;
  real_t acc = f.has(i, j, k) ? A.b * A.lam_out[o] : real_t(0);
  const int c[3] = {i, j, k};
  const long long step[3] = {1, A.sy, A.sz};
  const int lo[3] = {f.ilo, f.jlo, f.klo}, hi[3] = {f.ihi, f.jhi, f.khi};
  bool in_t[3] = {i >= f.ilo && i <= f.ihi, j >= f.jlo && j <= f.jhi, k >= f.klo && k <= f.khi};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const bool tr = in_t[(d + 1) % 3] && in_t[(d + 2) % 3];
    if (!tr) continue;
#pragma unroll
    for (int s = -3; s <= 3; s++) {
      const int src = c[d] - s;  // cell i = j - s e_d whose stencil offset s names j
      if (src >= lo[d] && src <= hi[d]) acc += A.q[(d * 7 + s + 3) * A.vol + o - s * step[d]];
    }
  }
  A.lam_us[o] += acc;
}

// Adjoint of the periodic fill: every ghost value is a copy of the interior
// cell one period away, so its adjoint is added there; ghosts are then zeroed.
__global__ void __launch_bounds__(256) fold_kernel(real_t* lam, Box gb, Box fb, int sy, int sz) {
  LSMHIP_CELL(fb);
  const int n[3] = {fb.nx(), fb.ny(), fb.nz()};
  const int c[3] = {i, j, k};
  const int glo[3] = {gb.ilo, gb.jlo, gb.klo}, ghi[3] = {gb.ihi, gb.jhi, gb.khi};
  real_t acc = real_t(0);
  for (int a = -1; a <= 1; a++)
    for (int bb = -1; bb <= 1; bb++)
      for (int cc = -1; cc <= 1; cc++) {
        if (a == 0 && bb == 0 && cc == 0) continue;
        const int q[3] = {c[0] + a * n[0], c[1] + bb * n[1], c[2] + cc * n[2]};
        bool ok = true;
        for (int d = 0; d < 3; d++) ok = ok && q[d] >= glo[d] && q[d] <= ghi[d];
        if (ok) acc += lam[(q[0] - gb.ilo) + (long long)sy * (q[1] - gb.jlo) + (long long)sz * (q[2] - gb.klo)];
      }
  lam[(i - gb.ilo) + (long long)sy * (j - gb.jlo) + (long long)sz * (k - gb.klo)] += acc;
}

__global__ void __launch_bounds__(256) zero_ghosts_kernel(real_t* lam, Box gb, Box fb, int sy, int sz) {
  LSMHIP_CELL(gb);
  if (!fb.has(i, j, k)) lam[(i - gb.ilo) + (long long)sy * (j - gb.jlo) + (long long)sz * (k - gb.klo)] = 0;
}

constexpr double kPi = 3.14159265358979323846;
constexpr double kOneOverPi = 0.31830988618379;

// Smoothed Heaviside of src/toolbox/lsm_calculus_toolbox.f:25-58 and its derivative.
__device__ __forceinline__ double heaviside(double v, double e) {
  if (v < -e) return 0.0;
  if (v > e) return 1.0;
  return 0.5 * (1.0 + v / e + kOneOverPi * sin(kPi * v / e));
}
__device__ __forceinline__ double heaviside_d(double v, double e) {
  if (v < -e || v > e) return 0.0;
  return 0.5 / e * (1.0 + kOneOverPi * kPi * cos(kPi * v / e));
}

struct MismatchOp {
  const real_t *phi, *tgt;
  Box gb;
  int sy, sz;
  double e, dV;
  __device__ double operator()(int i, int j, int k) const {
    const long long o = (i - gb.ilo) + (long long)sy * (j - gb.jlo) + (long long)sz * (k - gb.klo);
    const double r = heaviside((double)phi[o], e) - heaviside((double)tgt[o], e);
    return r * r * dV;
  }
};

__global__ void __launch_bounds__(256) mismatch_grad_kernel(real_t* grad, const real_t* phi, const real_t* tgt,
                                                            Box gb, Box fb, int sy, int sz, double e, double dV) {
  LSMHIP_CELL(gb);
  const long long o = (i - gb.ilo) + (long long)sy * (j - gb.jlo) + (long long)sz * (k - gb.klo);
  double g = 0.0;
  if (fb.has(i, j, k)) {
    const double p = (double)phi[o];
    g = 2.0 * (heaviside(p, e) - heaviside((double)tgt[o], e)) * heaviside_d(p, e) * dV;
  }
  grad[o] = (real_t)g;
}

}  // namespace

// ---------------------------------------------------------------------------
// Entry points (no counterpart in the reference).  Arrays are device arrays on
// the ghost box gb; the stage is the one lsmhip_weno5_stage_normal computes.
// ---------------------------------------------------------------------------

// Reverse of one normal-motion stage (stage 1, 2, 3 of TVD-RK3 or 22 = TVD-RK2
// stage 2).  Accumulates into lam_us (ghost box), lam_uc (fill box; ignored for
// stage 1, may be null) and lam_vn (fill box, only when vn_field is given; may
// be null).  lam_out is read on fb.  scratch: 21 * |gb| reals.  Returns 0 or -1.
LSMHIP_EXPORT int lsmhip_weno5_stage_normal_adjoint(real_t* lam_us, real_t* lam_uc, real_t* lam_vn,
                                                    const real_t* lam_out, const real_t* phi_stage,
                                                    LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), int stage, real_t vn,
                                                    const real_t* vn_field, real_t dt, real_t dx, real_t dy,
                                                    real_t dz, real_t* scratch) {
  LSMHIP_REQUIRE_ARRAYS_OR(-1, lam_us, lam_out, phi_stage, scratch);
  const char* who = "lsmhip_weno5_stage_normal_adjoint";
  AdjArgs A{};
  A.gb = LSMHIP_BOX(gb);
  A.fb = LSMHIP_BOX(fb);
  if (A.fb.empty()) return 0;
  if (!check_box(A.gb, A.fb.widen(3, 3, 3), who, "phi_stage") || !check_size(A.gb, who)) return -1;
  if (!lam_us || !lam_out || !phi_stage || !scratch) {
    set_error(kErrArg, "%s: lam_us, lam_out, phi_stage and scratch are required", who);
    return -1;
  }
  switch (stage) {
    case 1: A.a = 0; A.b = 1; break;
    case 2: A.a = real_t(0.75); A.b = real_t(0.25); break;
    case 3: A.a = real_t(1.0 / 3.0); A.b = real_t(2.0 / 3.0); break;
    case 22: A.a = real_t(0.5); A.b = real_t(0.5); break;
    default:
      set_error(kErrArg, "%s: stage must be 1, 2, 3 or 22, got %d", who, stage);
      return -1;
  }
  A.us = phi_stage;
  A.lam_out = lam_out;
  A.lam_us = lam_us;
  A.lam_uc = stage == 1 ? nullptr : lam_uc;
  A.lam_v = lam_vn;
  A.vnf = vn_field;
  A.vn = vn;
  A.dt = dt;
  A.inv[0] = real_t(1) / dx;
  A.inv[1] = real_t(1) / dy;
  A.inv[2] = real_t(1) / dz;
  A.q = scratch;
  A.vol = A.gb.volume();
  A.sy = A.gb.nx();
  A.sz = A.gb.nx() * A.gb.ny();
  hipLaunchKernelGGL(adj_weights_kernel, grid_for(A.fb), block2d(), 0, stream(), A);
  hipLaunchKernelGGL(adj_gather_kernel, grid_for(A.gb), block2d(), 0, stream(), A);
  return finish(who) ? 0 : -1;
}

// Adjoint of lsmhip_fill_ghosts(kind = periodic, all faces): adds every ghost
// entry of lam to the interior cell it copies and zeroes the ghosts.
LSMHIP_EXPORT int lsmhip_periodic_fill_adjoint(real_t* lam, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb)) {
  LSMHIP_REQUIRE_ARRAYS_OR(-1, lam);
  const char* who = "lsmhip_periodic_fill_adjoint";
  Box g = LSMHIP_BOX(gb), f = LSMHIP_BOX(fb);
  if (f.empty()) return 0;
  if (!check_box(g, f, who, "lam") || !check_size(g, who)) return -1;
  if (g.ilo < f.ilo - f.nx() || g.ihi > f.ihi + f.nx() || g.jlo < f.jlo - f.ny() || g.jhi > f.jhi + f.ny() ||
      g.klo < f.klo - f.nz() || g.khi > f.khi + f.nz()) {
    set_error(kErrArg, "%s: ghost layers wider than the period", who);
    return -1;
  }
  hipLaunchKernelGGL(fold_kernel, grid_for(f), block2d(), 0, stream(), lam, g, f, g.nx(), g.nx() * g.ny());
  hipLaunchKernelGGL(zero_ghosts_kernel, grid_for(g), block2d(), 0, stream(), lam, g, f, g.nx(), g.nx() * g.ny());
  return finish(who) ? 0 : -1;
}

// J = sum over fb of (H_eps(phi) - H_eps(target))^2 dx dy dz, accumulated in
// FP64 (H_eps of src/toolbox/lsm_calculus_toolbox.f:25-58), and, when grad is
// not null, dJ/dphi on gb (zero off fb).  Returns 0 or -1.
LSMHIP_EXPORT int lsmhip_heaviside_mismatch(double* loss, real_t* grad, const real_t* phi, const real_t* target,
                                            LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), double dx, double dy, double dz,
                                            double epsilon) {
  LSMHIP_REQUIRE_ARRAYS_OR(-1, loss, phi, target);
  const char* who = "lsmhip_heaviside_mismatch";
  Box g = LSMHIP_BOX(gb), f = LSMHIP_BOX(fb);
  if (!check_box(g, f, who, "phi") || !check_size(g, who)) return -1;
  MismatchOp op{phi, target, g, g.nx(), g.nx() * g.ny(), epsilon, dx * dy * dz};
  double s = 0.0;
  if (!f.empty() && !reduce::box_sum(who, f, op, &s)) return -1;
  if (loss) *loss = s;
  if (grad) {
    hipLaunchKernelGGL(mismatch_grad_kernel, grid_for(g), block2d(), 0, stream(), grad, phi, target, g, f, g.nx(),
                       g.nx() * g.ny(), epsilon, dx * dy * dz);
    if (!finish(who)) return -1;
  }
  return 0;
}
