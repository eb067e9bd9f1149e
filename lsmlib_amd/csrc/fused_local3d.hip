// fused_local3d.hip -- one TVD Runge-Kutta stage of normal motion on a narrow
// band in a single kernel, and the matching CFL step.
//
// The drop-in chain for a band stage runs HJ_WENO5_LOCAL (per axis: first
// differences over levels 0..3 into D1, then the plus/minus reconstruction at
// level 0), ZERO_OUT_LOCAL, ADD_[CONST_]NORMAL_VEL_TERM_LOCAL,
// MULTIPLY_CUT_OFF_LSE_RHS_LOCAL and TVD_RK3_STAGEn_LOCAL: nine launches that
// write and re-read six derivative arrays, D1 and the RHS.  Here one thread
// per level-0 point reads the 19-point cross of phi and the marks of the 18
// neighbours, forms the differences the drop-in would have stored (the 'big'
// value where a mark exceeds mark_D1), and goes straight to the RK update.
//
// Equality with the drop-in chain: the differences the chain reads at a
// level-0 point lie within three cells along an axis, and a cell that close to
// a level-0 point is in levels 0..3 of any band grown to at least three levels
// (first-claim growth reaches every cell within distance d by level d), so the
// chain has written them in this call.  A neighbour outside the band (mark 0)
// would be a stale difference in the chain; the kernel skips such a point and
// reports it (LSMHIP_ERR_INDEX) instead.  With that, FP64 results are bitwise
// the chain's; FP32 uses the rearranged WENO5 of the whole-grid fused stages.
#include <type_traits>

#include "local.hpp"
#include "stencil.hpp"

using namespace lsmhip;

namespace {

struct LocalStage {
  IndexList L;        // level-0 points
  Mask fbm;           // narrow_band, mark_fb
  signed char markD1;
  const real_t *us, *uc, *vnf;
  real_t* out;
  Box gb;             // us, uc, vnf and out share it
  int sy, sz;
  real_t inv[3];
  real_t vn, dt;
  int cutoff;
  real_t beta, gamma, c1, c2;
  int ncopy;          // entries lo..lo+ncopy-1 not updated by the stage take phi_n (0: none)
};

// The 19-point cross of phi around (i, j, k) (offset p) and the marks of the
// 18 neighbours whose differences the stencil reads.  cross_in_boxes: the
// cross lies in gb and the band's box (no memory read); load_cross issues the
// loads without testing them; cross_in_band: no neighbour lies outside the
// band (mark 0).
__device__ __forceinline__ bool cross_in_boxes(const LocalStage& a, int i, int j, int k) {
  const Box& nbb = a.fbm.b;
  return a.gb.has(i - 3, j - 3, k - 3) && a.gb.has(i + 3, j + 3, k + 3) && nbb.has(i - 2, j - 2, k - 2) &&
         nbb.has(i + 3, j + 3, k + 3);
}
__device__ __forceinline__ void load_cross(const LocalStage& a, int i, int j, int k, long long p, real_t f[3][7],
                                           signed char m[3][6]) {
  // addresses from the centre's pointers and uniform strides (x neighbours
  // become immediate offsets of one address)
  const real_t* c = a.us + p;
  const Box& b = a.fbm.b;
  const signed char* mc = (const signed char*)a.fbm.p + ((long long)(i - b.ilo) + (long long)a.fbm.sy * (j - b.jlo) +
                                                         (long long)a.fbm.sz * (k - b.klo));
  const long long st[3] = {1, a.sy, a.sz}, mst[3] = {1, a.fbm.sy, a.fbm.sz};
#pragma unroll
  for (int d = 0; d < 3; d++) {
#pragma unroll
    for (int s = -3; s <= 3; s++) f[d][s + 3] = c[s * st[d]];
#pragma unroll
    for (int s = -2; s <= 3; s++) m[d][s + 2] = mc[s * mst[d]];
  }
}
// Holds the loaded cross here: the compiler would otherwise sink the loads of
// phi below the tests of the marks (and those below the point's own), one
// dependent memory round trip each.  The empty asm takes every value, so all
// loads are issued (and waited for together) before any is tested.
__device__ __forceinline__ void hold(const real_t f[3][7], const signed char m[3][6]) {
#pragma unroll
  for (int d = 0; d < 3; d++) {
#pragma unroll
    for (int s = 0; s < 7; s++) asm volatile("" ::"v"(f[d][s]));
#pragma unroll
    for (int s = 0; s < 6; s++) asm volatile("" ::"v"((int)m[d][s]));
  }
}
__device__ __forceinline__ bool cross_in_band(const signed char m[3][6]) {
  bool out = false;
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int s = 0; s < 6; s++) out |= m[d][s] == 0;
  return !out;
}
// Plus and minus WENO5 derivatives along axis d from the gathered cross, with
// the differences HJ_WENO5_LOCAL stores (the 'big' value where a mark exceeds
// mark_D1).
__device__ __forceinline__ void plus_minus_local(const LocalStage& a, const real_t f[7], const signed char m[6],
                                                 int d, real_t& plus, real_t& minus) {
  real_t dd[6];  // dd[s+2] = D1(s), s = -2..3
#pragma unroll
  for (int s = -2; s <= 3; s++) dd[s + 2] = m[s + 2] <= a.markD1 ? f[s + 3] - f[s + 2] : kBig;
  const real_t inv = a.inv[d];
  if (sizeof(real_t) == 8) {
    plus = stencil::weno5(dd[5] * inv, dd[4] * inv, dd[3] * inv, dd[2] * inv, dd[1] * inv);
    minus = stencil::weno5(dd[0] * inv, dd[1] * inv, dd[2] * inv, dd[3] * inv, dd[4] * inv);
  } else {
    using namespace stencil::fastf;
    plus = (real_t)rev(win5((float)dd[1], (float)dd[2], (float)dd[3], (float)dd[4], (float)dd[5]), (float)dd[2],
                       (float)dd[3], (float)dd[4]) * inv;
    minus = (real_t)fwd(win5((float)dd[0], (float)dd[1], (float)dd[2], (float)dd[3], (float)dd[4]), (float)dd[1],
                        (float)dd[2], (float)dd[3]) * inv;
  }
}

__device__ __forceinline__ real_t local_sqrt(real_t x) {
  if (sizeof(real_t) == 8) return sqrt(x);
  return (real_t)__builtin_amdgcn_sqrtf((float)x);
}

// The RHS of one stage at a level-0 point: the Godunov normal-speed term
// (ADD_[CONST_]NORMAL_VEL_TERM_LOCAL on a zeroed RHS), times the cut-off of
// MULTIPLY_CUT_OFF_LSE_RHS_LOCAL when one is set, in the chain's operation order.
__device__ __forceinline__ real_t stage_rhs(const LocalStage& a, real_t v, const real_t g[6], real_t us) {
  real_t r = real_t(0);
  if (fabs(v) >= kZeroTol) r = r - v * local_sqrt(stencil::godunov_sq(v > real_t(0), g[0], g[1], g[2], g[3], g[4], g[5]));
  if (a.cutoff) {
    const real_t ab = fabs(us);
    real_t cut;
    if (ab <= a.beta) {
      cut = real_t(1);
    } else if (ab <= a.gamma) {
      const real_t q = ab - a.gamma;
      cut = (q * q * (real_t(2) * ab + a.c1)) / a.c2;
    } else {
      cut = real_t(0);
    }
    r = cut * r;
  }
  return r;
}

// Stored in rhs[t] where the stencil may not be evaluated: a signalling-NaN
// bit pattern, which no arithmetic produces (the hardware's NaNs are quiet),
// written and compared as an integer, never as a real.
typedef std::conditional<sizeof(real_t) == 8, unsigned long long, unsigned int>::type rhs_bits_t;
constexpr rhs_bits_t kSkipBits = sizeof(real_t) == 8 ? (rhs_bits_t)0x7ff40000000051abULL : (rhs_bits_t)0x7fa051abU;

template <int STAGE>
__device__ __forceinline__ real_t rk_combine(real_t uc, real_t us, real_t dt, real_t r) {
  if (STAGE == 1) return uc + dt * r;
  if (STAGE == 2) return real_t(0.75) * uc + real_t(0.25) * (us + dt * r);
  if (STAGE == 3) return real_t(1.0 / 3.0) * uc + real_t(2.0 / 3.0) * (us + dt * r);
  return real_t(0.5) * (uc + us + dt * r);  // TVD-RK2 stage 2
}

// Entry t that the stage does not update, within the copy range: out = phi_n
// there (phi_stage at stage 1, where it is phi_n), as COPY_DATA_LOCAL(out, phi_n)
// over those levels before the stage would leave it.
template <int STAGE>
__device__ __forceinline__ void pass_through(const LocalStage& a, long long t, int* flag) {
  if (t >= a.ncopy) return;
  int i, j, k;
  a.L.at(a.L.lo + (int)t, i, j, k);
  if (!a.gb.has(i, j, k)) {
    *flag = 1;
    return;
  }
  const long long p = (long long)(i - a.gb.ilo) + (long long)a.sy * (j - a.gb.jlo) + (long long)a.sz * (k - a.gb.klo);
  a.out[p] = STAGE == 1 ? a.us[p] : a.uc[p];
}

template <int STAGE>
__global__ void __launch_bounds__(kT) local_normal_stage_kernel(LocalStage a, int* flag) {
  const long long t = xcd_entry();
  if (t > (long long)a.L.hi - a.L.lo) {
    pass_through<STAGE>(a, t, flag);
    return;
  }
  int i, j, k;
  a.L.at(a.L.lo + (int)t, i, j, k);
  if (!a.fbm.b.has(i, j, k) || !a.gb.has(i, j, k)) {
    *flag = 1;
    return;
  }
  const long long p = (long long)(i - a.gb.ilo) + (long long)a.sy * (j - a.gb.jlo) + (long long)a.sz * (k - a.gb.klo);
  if (!cross_in_boxes(a, i, j, k)) {  // an error unless the stage leaves the point
    if (a.fbm.at(i, j, k) > a.fbm.mark)
      pass_through<STAGE>(a, t, flag);
    else
      *flag = 1;
    return;
  }
  // The point's own mark, its speed and phi_n, and the cross go out together
  // (one memory round trip after the index list, not two); the cross of a
  // point the stage does not update is loaded but never tested.
  const signed char mark = a.fbm.at(i, j, k);
  const real_t v = a.vnf ? a.vnf[p] : a.vn;
  const real_t ucv = STAGE == 1 ? real_t(0) : a.uc[p];
  real_t f[3][7];
  signed char m[3][6];
  load_cross(a, i, j, k, p, f, m);
  hold(f, m);
  asm volatile("" ::"v"((int)mark));
  if (mark > a.fbm.mark) {  // mark above mark_fb
    pass_through<STAGE>(a, t, flag);
    return;
  }
  if (!cross_in_band(m)) {
    *flag = 1;
    return;
  }
  real_t g[6];  // xp, yp, zp, xm, ym, zm
  plus_minus_local(a, f[0], m[0], 0, g[0], g[3]);
  plus_minus_local(a, f[1], m[1], 1, g[1], g[4]);
  plus_minus_local(a, f[2], m[2], 2, g[2], g[5]);
  const real_t us = f[0][3];
  a.out[p] = rk_combine<STAGE>(STAGE == 1 ? us : ucv, us, a.dt, stage_rhs(a, v, g, us));
}

// Stage from a stored RHS (lsmhip_weno5_local_normal_dt_rhs): the same points,
// copies and RK combination as local_normal_stage_kernel, with rhs[t] in place
// of the derivatives; an entry holding the skip pattern is skipped and flagged.
template <int STAGE>
__global__ void __launch_bounds__(kT) local_stage_from_rhs_kernel(LocalStage a, const real_t* rhs, int* flag) {
  const long long t = (long long)blockIdx.x * kT + threadIdx.x;
  if (t > (long long)a.L.hi - a.L.lo) {
    pass_through<STAGE>(a, t, flag);
    return;
  }
  const rhs_bits_t rb = reinterpret_cast<const rhs_bits_t*>(rhs)[t];  // with the index
  int i, j, k;
  a.L.at(a.L.lo + (int)t, i, j, k);
  if (!a.fbm.b.has(i, j, k) || !a.gb.has(i, j, k)) {
    *flag = 1;
    return;
  }
  const long long p = (long long)(i - a.gb.ilo) + (long long)a.sy * (j - a.gb.jlo) + (long long)a.sz * (k - a.gb.klo);
  // the mark and the point's values together, before either is tested
  const signed char mark = a.fbm.at(i, j, k);
  const real_t us = a.us[p];
  const real_t ucv = STAGE == 1 ? real_t(0) : a.uc[p];
  asm volatile("" ::"v"((int)mark), "v"(us), "v"(ucv));
  if (mark > a.fbm.mark) {  // mark above mark_fb
    pass_through<STAGE>(a, t, flag);
    return;
  }
  if (rb == kSkipBits) {
    *flag = 1;
    return;
  }
  real_t r;
  __builtin_memcpy(&r, &rb, sizeof(r));
  a.out[p] = rk_combine<STAGE>(STAGE == 1 ? us : ucv, us, a.dt, r);
}

// CFL of normal motion at one level-0 point: lsm3dComputeStableNormalVelDtLOCAL
// with a speed field, the whole-grid constant-speed formula otherwise.
// With rhs set it also stores the stage RHS of the point at rhs[t] (the skip
// pattern where the stencil may not be evaluated), for a first stage from
// lsmhip_local_stage_from_rhs.
struct LocalDtOp {
  LocalStage a;
  real_t mdx2;
  int* flag;
  real_t* rhs;
  // The point's mark (M is a.fbm), its speed and the cross are loaded
  // together; a point above mark_fb gives no value and no error.
  __device__ bool marked(long long t, int i, int j, int k, const Mask& M, double& r) const {
    if (!a.gb.has(i, j, k) || !cross_in_boxes(a, i, j, k)) {
      if (M.at(i, j, k) > M.mark) return true;
      if (rhs) reinterpret_cast<rhs_bits_t*>(rhs)[t] = kSkipBits;
      if (a.gb.has(i, j, k)) *flag = 1;  // the cross leaves gb or the band's box
      return false;
    }
    const long long p =
        (long long)(i - a.gb.ilo) + (long long)a.sy * (j - a.gb.jlo) + (long long)a.sz * (k - a.gb.klo);
    const signed char mark = M.at(i, j, k);
    const real_t v = a.vnf ? a.vnf[p] : a.vn;
    real_t f[3][7];
    signed char m[3][6];
    load_cross(a, i, j, k, p, f, m);
    hold(f, m);
    asm volatile("" ::"v"((int)mark), "v"(v));
    if (mark > M.mark) return true;
    if (!cross_in_band(m)) {
      *flag = 1;
      if (rhs) reinterpret_cast<rhs_bits_t*>(rhs)[t] = kSkipBits;
      return false;
    }
    real_t g[6];  // xp, yp, zp, xm, ym, zm
    plus_minus_local(a, f[0], m[0], 0, g[0], g[3]);
    plus_minus_local(a, f[1], m[1], 1, g[1], g[4]);
    plus_minus_local(a, f[2], m[2], 2, g[2], g[5]);
    if (rhs) rhs[t] = stage_rhs(a, v, g, f[0][3]);
    const real_t av = fabs(v);
    real_t n;
    const real_t h = stencil::normal_h_over_dx(av, g[0], g[1], g[2], g[3], g[4], g[5], mdx2, a.inv[0], a.inv[1],
                                               a.inv[2], &n);
    // the variable-speed routine admits only points with |grad phi| > zero_tol
    // (lsm_math_utils3d_local.f.in:405-416); the constant-speed one all points
    r = (a.vnf && !(n > kZeroTol)) ? -1.0 : (double)h;
    return true;
  }
};

LocalStage make_stage(real_t* out, const real_t* us, const real_t* uc, const Box& gb, const int* ix, const int* iy,
                      const int* iz, int nlo, int nhi, const unsigned char* nb, const Box& nbb, unsigned char mark_fb,
                      unsigned char mark_d1, real_t dx, real_t dy, real_t dz, real_t vn, const real_t* vnf) {
  LocalStage a{};
  a.L = IndexList{ix, iy, iz, nlo, nhi, nlo};
  a.fbm = make_mask(nb, nbb, &mark_fb, 1);
  a.markD1 = (signed char)mark_d1;
  a.us = us;
  a.uc = uc;
  a.vnf = vnf;
  a.out = out;
  a.gb = gb;
  a.sy = gb.nx();
  a.sz = gb.nx() * gb.ny();
  a.inv[0] = real_t(1.0) / dx;
  a.inv[1] = real_t(1.0) / dy;
  a.inv[2] = real_t(1.0) / dz;
  a.vn = vn;
  return a;
}

// The cut-off of MULTIPLY_CUT_OFF_LSE_RHS_LOCAL (none when gamma <= 0).
void set_cutoff(LocalStage& a, real_t beta, real_t gamma) {
  a.cutoff = gamma > real_t(0);
  a.beta = beta;
  a.gamma = gamma;
  a.c1 = gamma - real_t(3) * beta;
  const real_t c2 = gamma - beta;
  a.c2 = c2 * c2 * c2;
}

bool check_stage(const char* who, int stage) {
  if (stage == 1 || stage == 2 || stage == 3 || stage == 22) return true;
  set_error(kErrArg, "%s: stage %d is not 1, 2, 3 or 22", who, stage);
  return false;
}

}  // namespace

LSMHIP_EXPORT int lsmhip_weno5_local_stage_normal(LSMLIB_REAL* phi_out, const LSMLIB_REAL* phi_stage,
                                                  const LSMLIB_REAL* phi_n, LSMHIP_BOX_ARGS(gb), int stage,
                                                  LSMLIB_REAL vn, const LSMLIB_REAL* vn_field, LSMLIB_REAL dt,
                                                  LSMLIB_REAL dx, LSMLIB_REAL dy, LSMLIB_REAL dz, const int* index_x,
                                                  const int* index_y, const int* index_z, int nlo_index0,
                                                  int nhi_index0, int nhi_copy, const unsigned char* narrow_band,
                                                  LSMHIP_BOX_ARGS(nb_gb), unsigned char mark_fb,
                                                  unsigned char mark_D1, LSMLIB_REAL beta, LSMLIB_REAL gamma) {
  const char* who = "lsmhip_weno5_local_stage_normal";
  const Box gb = box_of(ilo_gb, ihi_gb, jlo_gb, jhi_gb, klo_gb, khi_gb);
  const Box nbb = box_of(ilo_nb_gb, ihi_nb_gb, jlo_nb_gb, jhi_nb_gb, klo_nb_gb, khi_nb_gb);
  if (!check_stage(who, stage)) return -1;
  if (stage != 1 && !phi_n) {
    set_error(kErrArg, "%s: stage %d needs phi_n", who, stage);
    return -1;
  }
  if (!phi_out || !phi_stage || !index_x || !index_y || !index_z || !narrow_band) {
    set_error(kErrArg, "%s: a required array is NULL", who);
    return -1;
  }
  if (!check_size(gb, who) || !check_size(nbb, who)) return -1;
  const int hi = nhi_copy > nhi_index0 ? nhi_copy : nhi_index0;
  if (hi < nlo_index0) return 0;
  LocalStage a = make_stage(phi_out, phi_stage, phi_n, gb, index_x, index_y, index_z, nlo_index0, nhi_index0,
                            narrow_band, nbb, mark_fb, mark_D1, dx, dy, dz, vn, vn_field);
  a.ncopy = nhi_copy >= nlo_index0 ? nhi_copy - nlo_index0 + 1 : 0;
  a.dt = dt;
  set_cutoff(a, beta, gamma);
  int* flag = device_flag();
  if (!flag) return -1;
  const dim3 grid(xcd_grid((long long)hi - nlo_index0 + 1));
  switch (stage) {
    case 1: hipLaunchKernelGGL(local_normal_stage_kernel<1>, grid, dim3(kT), 0, stream(), a, flag); break;
    case 2: hipLaunchKernelGGL(local_normal_stage_kernel<2>, grid, dim3(kT), 0, stream(), a, flag); break;
    case 3: hipLaunchKernelGGL(local_normal_stage_kernel<3>, grid, dim3(kT), 0, stream(), a, flag); break;
    default: hipLaunchKernelGGL(local_normal_stage_kernel<22>, grid, dim3(kT), 0, stream(), a, flag); break;
  }
  if (!check_hip(hipGetLastError(), who)) return -1;
  if (!finish(who)) return -1;
  if (synchronous()) take_device_flag(who);
  return lsmhip_get_last_error() == 0 ? 0 : -1;
}

static int local_normal_dt(const char* who, real_t* dt, real_t* rhs, const real_t* phi, const Box& gb, real_t vn,
                           const real_t* vn_field, real_t dx, real_t dy, real_t dz, real_t cfl_number,
                           const int* index_x, const int* index_y, const int* index_z, int nlo_index0,
                           int nhi_index0, const unsigned char* narrow_band, const Box& nbb, unsigned char mark_fb,
                           unsigned char mark_D1, real_t beta, real_t gamma) {
  if (!dt || !phi || !index_x || !index_y || !index_z || !narrow_band) {
    set_error(kErrArg, "%s: a required array is NULL", who);
    return -1;
  }
  if (!check_size(gb, who) || !check_size(nbb, who)) return -1;
  int* flag = device_flag();
  if (!flag) return -1;
  LocalStage a = make_stage(nullptr, phi, nullptr, gb, index_x, index_y, index_z, nlo_index0, nhi_index0,
                            narrow_band, nbb, mark_fb, mark_D1, dx, dy, dz, vn, vn_field);
  set_cutoff(a, beta, gamma);
  real_t h = dx > dy ? dx : dy;
  h = h > dz ? h : dz;
  double m = -1.0;
  bool any = false;
  if (!points_max(who, a.L, a.fbm, LocalDtOp{a, h * h, flag, rhs}, &m, &any)) return -1;
  if (!any) m = -1.0;
  *dt = cfl_number / (real_t(m) + kTiny);
  return lsmhip_get_last_error() == 0 ? 0 : -1;
}

LSMHIP_EXPORT int lsmhip_weno5_local_normal_dt(LSMLIB_REAL* dt, const LSMLIB_REAL* phi, LSMHIP_BOX_ARGS(gb),
                                               LSMLIB_REAL vn, const LSMLIB_REAL* vn_field, LSMLIB_REAL dx,
                                               LSMLIB_REAL dy, LSMLIB_REAL dz, LSMLIB_REAL cfl_number,
                                               const int* index_x, const int* index_y, const int* index_z,
                                               int nlo_index0, int nhi_index0, const unsigned char* narrow_band,
                                               LSMHIP_BOX_ARGS(nb_gb), unsigned char mark_fb, unsigned char mark_D1) {
  return local_normal_dt("lsmhip_weno5_local_normal_dt", dt, nullptr, phi,
                         box_of(ilo_gb, ihi_gb, jlo_gb, jhi_gb, klo_gb, khi_gb), vn, vn_field, dx, dy, dz, cfl_number,
                         index_x, index_y, index_z, nlo_index0, nhi_index0, narrow_band,
                         box_of(ilo_nb_gb, ihi_nb_gb, jlo_nb_gb, jhi_nb_gb, klo_nb_gb, khi_nb_gb), mark_fb, mark_D1,
                         real_t(0), real_t(0));
}

LSMHIP_EXPORT int lsmhip_weno5_local_normal_dt_rhs(LSMLIB_REAL* dt, LSMLIB_REAL* rhs, const LSMLIB_REAL* phi,
                                                   LSMHIP_BOX_ARGS(gb), LSMLIB_REAL vn, const LSMLIB_REAL* vn_field,
                                                   LSMLIB_REAL dx, LSMLIB_REAL dy, LSMLIB_REAL dz,
                                                   LSMLIB_REAL cfl_number, const int* index_x, const int* index_y,
                                                   const int* index_z, int nlo_index0, int nhi_index0,
                                                   const unsigned char* narrow_band, LSMHIP_BOX_ARGS(nb_gb),
                                                   unsigned char mark_fb, unsigned char mark_D1, LSMLIB_REAL beta,
                                                   LSMLIB_REAL gamma) {
  const char* who = "lsmhip_weno5_local_normal_dt_rhs";
  if (!rhs) {
    set_error(kErrArg, "%s: a required array is NULL", who);
    return -1;
  }
  return local_normal_dt(who, dt, rhs, phi, box_of(ilo_gb, ihi_gb, jlo_gb, jhi_gb, klo_gb, khi_gb), vn, vn_field,
                         dx, dy, dz, cfl_number, index_x, index_y, index_z, nlo_index0, nhi_index0, narrow_band,
                         box_of(ilo_nb_gb, ihi_nb_gb, jlo_nb_gb, jhi_nb_gb, klo_nb_gb, khi_nb_gb), mark_fb, mark_D1,
                         beta, gamma);
}

LSMHIP_EXPORT int lsmhip_local_stage_from_rhs(LSMLIB_REAL* phi_out, const LSMLIB_REAL* phi_stage,
                                              const LSMLIB_REAL* phi_n, LSMHIP_BOX_ARGS(gb), int stage,
                                              const LSMLIB_REAL* rhs, LSMLIB_REAL dt, const int* index_x,
                                              const int* index_y, const int* index_z, int nlo_index0,
                                              int nhi_index0, int nhi_copy, const unsigned char* narrow_band,
                                              LSMHIP_BOX_ARGS(nb_gb), unsigned char mark_fb) {
  const char* who = "lsmhip_local_stage_from_rhs";
  const Box gb = box_of(ilo_gb, ihi_gb, jlo_gb, jhi_gb, klo_gb, khi_gb);
  const Box nbb = box_of(ilo_nb_gb, ihi_nb_gb, jlo_nb_gb, jhi_nb_gb, klo_nb_gb, khi_nb_gb);
  if (!check_stage(who, stage)) return -1;
  if (stage != 1 && !phi_n) {
    set_error(kErrArg, "%s: stage %d needs phi_n", who, stage);
    return -1;
  }
  if (!phi_out || !phi_stage || !rhs || !index_x || !index_y || !index_z || !narrow_band) {
    set_error(kErrArg, "%s: a required array is NULL", who);
    return -1;
  }
  if (!check_size(gb, who) || !check_size(nbb, who)) return -1;
  const int hi = nhi_copy > nhi_index0 ? nhi_copy : nhi_index0;
  if (hi < nlo_index0) return 0;
  LocalStage a = make_stage(phi_out, phi_stage, phi_n, gb, index_x, index_y, index_z, nlo_index0, nhi_index0,
                            narrow_band, nbb, mark_fb, mark_fb, real_t(1), real_t(1), real_t(1), real_t(0), nullptr);
  a.ncopy = nhi_copy >= nlo_index0 ? nhi_copy - nlo_index0 + 1 : 0;
  a.dt = dt;
  int* flag = device_flag();
  if (!flag) return -1;
  const long long n = (long long)hi - nlo_index0 + 1;
  const dim3 grid((unsigned)((n + kT - 1) / kT));
  switch (stage) {
    case 1: hipLaunchKernelGGL(local_stage_from_rhs_kernel<1>, grid, dim3(kT), 0, stream(), a, rhs, flag); break;
    case 2: hipLaunchKernelGGL(local_stage_from_rhs_kernel<2>, grid, dim3(kT), 0, stream(), a, rhs, flag); break;
    case 3: hipLaunchKernelGGL(local_stage_from_rhs_kernel<3>, grid, dim3(kT), 0, stream(), a, rhs, flag); break;
    default: hipLaunchKernelGGL(local_stage_from_rhs_kernel<22>, grid, dim3(kT), 0, stream(), a, rhs, flag); break;
  }
  if (!check_hip(hipGetLastError(), who)) return -1;
  if (!finish(who)) return -1;
  if (synchronous()) take_device_flag(who);
  return lsmhip_get_last_error() == 0 ? 0 : -1;
}
