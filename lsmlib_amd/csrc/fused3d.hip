// fused3d.hip -- one TVD Runge-Kutta stage of the level-set equation in a
// single kernel: HJ-WENO5 one-sided derivatives, the Hamiltonian (normal
// motion with Godunov selection, velocity-upwinded advection, or the
// reinitialization RHS) and the RK combination, with no derivative, RHS or
// velocity array written to memory.
//
// Memory layout and traffic: phi_stage, phi_n and the output share one ghost
// box.  A block owns a TX x TY tile of the (i,j) plane and marches along k
// through a chunk of KC planes.  Each thread keeps the seven values of its own
// column around plane k (the WENO5 reach is +-3) in registers (in the
// advection and normal-motion kernel with the next plane's value in flight),
// so the z stencil needs neither LDS nor index arithmetic; the current plane
// with its 3-cell x/y halo is staged in LDS for the x and y stencils
// (interior from the threads' windows, the halo loaded by one thread per halo
// cell).  Each plane is read from global memory once per block;
// neighbouring tiles share halos through L2: blocks are numbered so that each
// XCD (blocks are dealt round-robin to the 8 XCDs) processes one contiguous
// range of tiles.
//
// The per-cell arithmetic is that of the drop-in chain (stencil.hpp), so in
// the FP64 build a stage is bitwise equal to
//   HJ_WENO5 / UPWIND_HJ_WENO5 -> zero RHS -> add term -> TVD_RK3_STAGEn
// of the reference on the same inputs.
#include "lsmhip_core.hpp"
#include "stencil.hpp"

using namespace lsmhip;

namespace {

constexpr int TX = 64, TY = 8, H = 3;
constexpr int PX = TX + 2 * H, PY = TY + 2 * H, PLANE = PX * PY;

enum RhsKind { kNormalConst = 0, kNormalField = 1, kAdvectField = 2, kAdvectDeform = 3, kReinit = 4 };

struct FusedArgs {
  const real_t* us;  // field differentiated this stage (ghosts filled)
  const real_t* uc;  // phi at the start of the step (stages 2 and 3)
  real_t* out;
  Box gb, fb;
  int sy, sz;
  real_t ix, iy, iz, dt;
  int rhs;
  real_t vn;                          // constant normal speed
  const real_t *vnf, *vx, *vy, *vz;   // per-cell fields (gb addressing)
  const real_t* phi0;                 // reinitialization sign source
  int use_phi0;
  real_t dx_sq;                       // max(dx,dy,dz)^2 for the smoothed sign
  const real_t *ta[3], *tb[3];        // deformation tables A = sin^2(pi x), B = sin(2 pi x), per axis
  real_t tfac;                        // cos(pi t / T) of the deformation field
  int nti, ntj, ntk, kc;              // tiles in i, j, chunks in k, planes per chunk
};

// Deformation velocity u = 2 A(x)B(y)B(z)c, v = -B(x)A(y)B(z)c, w = -B(x)B(y)A(z)c
// from per-axis tables indexed from the ghost-box origin.
__device__ __forceinline__ void deform_velocity(const real_t* const ta[3], const real_t* const tb[3], int ii, int jj,
                                                int kk, real_t c, real_t& u, real_t& v, real_t& w) {
  u = real_t(2) * ta[0][ii] * tb[1][jj] * tb[2][kk] * c;
  v = -(tb[0][ii] * ta[1][jj] * tb[2][kk] * c);
  w = -(tb[0][ii] * tb[1][jj] * ta[2][kk] * c);
}

// WENO5 derivatives of the fused stage: the reference's operation order in
// FP64 (bitwise equal to the drop-in chain), the rearranged FP32 forms in FP32.
__device__ __forceinline__ void weno_pm(const double f[7], double inv, double& p, double& m) {
  stencil::plus_minus<5>(f, inv, p, m);
}
__device__ __forceinline__ void weno_pm(const float f[7], float inv, float& p, float& m) {
  stencil::fastf::plus_minus(f, inv, p, m);
}
__device__ __forceinline__ double weno_up(const double f[7], double inv, bool vpos) {
  return stencil::upwind<5>(f, inv, vpos);
}
__device__ __forceinline__ float weno_up(const float f[7], float inv, bool vpos) {
  return stencil::fastf::upwind_wave(f, inv, vpos);
}

// Square root and reinitialization RHS of the fused stages: the reference's
// operations in FP64; the hardware square root (about 1 ulp) in FP32.
__device__ __forceinline__ double stage_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float stage_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double stage_reinit(double s, bool u0, double xp, double yp, double zp, double xm,
                                               double ym, double zm, double dx_sq) {
  return stencil::reinit_rhs(s, u0, xp, yp, zp, xm, ym, zm, dx_sq);
}
__device__ __forceinline__ float stage_reinit(float s, bool u0, float xp, float yp, float zp, float xm, float ym,
                                              float zm, float dx_sq) {
  return stencil::fastf::reinit_rhs(s, u0, xp, yp, zp, xm, ym, zm, dx_sq);
}

template <int STAGE>
__device__ __forceinline__ real_t rk_combine(real_t uc, real_t us, real_t dt, real_t r) {
  if (STAGE == 1) return uc + dt * r;
  if (STAGE == 2) return real_t(0.75) * uc + real_t(0.25) * (us + dt * r);
  if (STAGE == 3) return real_t(1.0 / 3.0) * uc + real_t(2.0 / 3.0) * (us + dt * r);
  return real_t(0.5) * (uc + us + dt * r);  // TVD-RK2 stage 2
}

// Position (px, py) in the PX x PY plane tile of halo element h (0 <= h < NHALO):
// three full rows above, three below, then three columns left and right of
// the TY interior rows.
constexpr int NHALO = PLANE - TX * TY;
__device__ __forceinline__ void halo_pos(int h, int& px, int& py) {
  if (h < H * PX) {
    py = h / PX;
    px = h % PX;
  } else if (h < 2 * H * PX) {
    py = H + TY + (h - H * PX) / PX;
    px = (h - H * PX) % PX;
  } else {
    const int t = h - 2 * H * PX, c = t % (2 * H);
    py = H + t / (2 * H);
    px = c < H ? c : TX + c;
  }
}

template <int RHS, int STAGE>
__global__ void __launch_bounds__(TX* TY) fused_stage_kernel(FusedArgs a) {
  __shared__ real_t plane[PLANE];
  const int tid = threadIdx.y * TX + threadIdx.x;

  // XCD-aware tile order: block b runs on XCD b % 8; give each XCD a contiguous tile range.
  const int nt = a.nti * a.ntj * a.ntk;
  const int b = blockIdx.x;
  const int per = (nt + 7) / 8;
  const int t = (b % 8) * per + b / 8;
  if (t >= nt) return;
  const int ti = t % a.nti, tj = (t / a.nti) % a.ntj, tk = t / (a.nti * a.ntj);
  const int i0 = a.fb.ilo + ti * TX, j0 = a.fb.jlo + tj * TY;
  const int kb = a.fb.klo + tk * a.kc;
  const int ke = min(kb + a.kc - 1, a.fb.khi);
  const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
  const Box gb = a.gb;
  const bool active = i <= a.fb.ihi && j <= a.fb.jhi;
  // own column: needed by neighbours whenever it lies in the ghost box
  const bool own_in = i <= gb.ihi && j <= gb.jhi;
  const long long own = (long long)(i - gb.ilo) + (long long)a.sy * (j - gb.jlo);
  // halo element of this thread (the same slot in every plane), -1 when none or outside the ghost box
  long long hoff = -1;
  int hslot = 0;
  if (tid < NHALO) {
    int px, py;
    halo_pos(tid, px, py);
    const int gi = i0 - H + px, gj = j0 - H + py;
    hslot = py * PX + px;
    if (gi >= gb.ilo && gi <= gb.ihi && gj >= gb.jlo && gj <= gb.jhi)
      hoff = (long long)(gi - gb.ilo) + (long long)a.sy * (gj - gb.jlo);
  }
  const int lx = threadIdx.x + H, ly = threadIdx.y + H;
  const real_t* us = a.us;
  const long long sz = a.sz;

  real_t ta0 = 0, tb0 = 0, ta1 = 0, tb1 = 0;
  if (RHS == kAdvectDeform && active) {
    ta0 = a.ta[0][i - gb.ilo];
    tb0 = a.tb[0][i - gb.ilo];
    ta1 = a.ta[1][j - gb.jlo];
    tb1 = a.tb[1][j - gb.jlo];
  }
  // FP32: the deformation velocity divided by the grid spacing, u/dx = X0 B(z),
  // v/dy = X1 B(z), w/dz = X2 A(z), with the k-independent factors formed once
  const real_t X0 = real_t(2) * a.tfac * a.ix * ta0 * tb1, X1 = -(a.tfac * a.iy) * tb0 * ta1,
               X2 = -(a.tfac * a.iz) * tb0 * tb1;

  // plane k + d of the own column
  auto zload = [&](int k, int d) { return own_in ? us[own + sz * (k + d - gb.klo)] : real_t(0); };
  // the halo element and the per-cell inputs of plane k
  struct In {
    real_t hv, uc, c0, c1, c2;
  };
  auto load_in = [&](int k) {
    const long long zoff = sz * (k - gb.klo);
    const long long off = own + zoff;
    In v{real_t(0), real_t(0), real_t(0), real_t(0), real_t(0)};
    if (hoff >= 0) v.hv = us[hoff + zoff];
    if (active) {
      if (STAGE != 1) v.uc = a.uc[off];
      if (RHS == kAdvectField) {
        v.c0 = a.vx[off];
        v.c1 = a.vy[off];
        v.c2 = a.vz[off];
      } else if (RHS == kReinit) {
        v.c0 = a.use_phi0 ? a.phi0[off] : real_t(0);
      } else if (RHS == kNormalField) {
        v.c0 = a.vnf[off];
      }
    }
    return v;
  };
  // cell (i, j, k) with its z window w[s + 3] = u(i, j, k + s) and inputs in
  auto cell = [&](int k, const real_t (&w)[7], const In& in) {
    const long long off = own + sz * (k - gb.klo);
    const real_t hv = in.hv, ucv = in.uc, c0 = in.c0, c1 = in.c1, c2 = in.c2;
    __syncthreads();  // the previous plane's reads are done
    plane[ly * PX + lx] = w[3];
    if (tid < NHALO) plane[hslot] = hv;
    __syncthreads();
    if (active) {
      real_t fx[7], fy[7], fz[7];
#pragma unroll
      for (int s = -H; s <= H; s++) {
        fx[s + 3] = plane[ly * PX + lx + s];
        fy[s + 3] = plane[(ly + s) * PX + lx];
        fz[s + 3] = w[s + 3];
      }
      real_t r;
      if (sizeof(real_t) == 4 && RHS == kAdvectDeform) {
        // r = -(u phi_x + v phi_y + w phi_z) with the derivatives' 1/dx folded into the velocity;
        // |u| < zero_tol  <=>  |u/dx| < zero_tol/dx
        const real_t ux = X0 * a.tb[2][k - gb.klo], vy = X1 * a.tb[2][k - gb.klo], wz = X2 * a.ta[2][k - gb.klo];
        const real_t px = fabs(ux) < kZeroTol * a.ix ? real_t(0) : weno_up(fx, real_t(1), ux > real_t(0));
        const real_t py = fabs(vy) < kZeroTol * a.iy ? real_t(0) : weno_up(fy, real_t(1), vy > real_t(0));
        const real_t pz = fabs(wz) < kZeroTol * a.iz ? real_t(0) : weno_up(fz, real_t(1), wz > real_t(0));
        r = real_t(0) - (ux * px + vy * py + wz * pz);
      } else if (RHS == kAdvectField || RHS == kAdvectDeform) {
        real_t u, v, vw;
        if (RHS == kAdvectField) {
          u = c0;
          v = c1;
          vw = c2;
        } else {
          const real_t ta2 = a.ta[2][k - gb.klo], tb2 = a.tb[2][k - gb.klo], c = a.tfac;
          u = real_t(2) * ta0 * tb1 * tb2 * c;
          v = -(tb0 * ta1 * tb2 * c);
          vw = -(tb0 * tb1 * ta2 * c);
        }
        const real_t px = fabs(u) < kZeroTol ? real_t(0) : weno_up(fx, a.ix, u > real_t(0));
        const real_t py = fabs(v) < kZeroTol ? real_t(0) : weno_up(fy, a.iy, v > real_t(0));
        const real_t pz = fabs(vw) < kZeroTol ? real_t(0) : weno_up(fz, a.iz, vw > real_t(0));
        r = real_t(0) - (u * px + v * py + vw * pz);
      } else {
        real_t xp, xm, yp, ym, zp, zm;
        weno_pm(fx, a.ix, xp, xm);
        weno_pm(fy, a.iy, yp, ym);
        weno_pm(fz, a.iz, zp, zm);
        if (RHS == kReinit) {
          const real_t s = a.use_phi0 ? c0 : fx[3];
          r = stage_reinit(s, a.use_phi0 != 0, xp, yp, zp, xm, ym, zm, a.dx_sq);
        } else {
          const real_t v = RHS == kNormalField ? c0 : a.vn;
          r = real_t(0);
          if (fabs(v) >= kZeroTol) r = r - v * stage_sqrt(stencil::godunov_sq(v > real_t(0), xp, yp, zp, xm, ym, zm));
        }
      }
      const real_t uc = STAGE == 1 ? fx[3] : ucv;
      a.out[off] = rk_combine<STAGE>(uc, fx[3], a.dt, r);
    }
  };

  // The z window, one plane ahead: plane k + 4 of the own column (and the
  // halo and inputs of plane k + 1) load while plane k is worked on.  FP32
  // keeps the window in a ring of eight registers: in a group of eight planes
  // starting at k, plane k + q has u(i, j, k + q + s) in z[(q + s + 3) % 8]
  // for s = -3 .. 4, and the group is unrolled, so the window advances
  // without moving registers.  The last planes of the chunk (fewer than
  // eight), and every plane in FP64 (where the unrolled group takes registers
  // enough to cost a resident block per CU), shift the window instead.
  real_t z[8];
#pragma unroll
  for (int s = 0; s < 7; s++) z[s] = zload(kb, s - H);
  In cur = load_in(kb);
  int k = kb;
  if constexpr (sizeof(real_t) == 4) {
    for (; k + 7 <= ke; k += 8) {
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const bool more = k + q < ke;
        if (more) z[(q + 7) % 8] = zload(k + q, H + 1);
        const In nxt = more ? load_in(k + q + 1) : In{};
        const real_t w[7] = {z[q % 8],       z[(q + 1) % 8], z[(q + 2) % 8], z[(q + 3) % 8],
                             z[(q + 4) % 8], z[(q + 5) % 8], z[(q + 6) % 8]};
        cell(k + q, w, cur);
        cur = nxt;
      }
    }
  }
  for (; k <= ke; k++) {
    const bool more = k < ke;
    if (more) z[7] = zload(k, H + 1);
    const In nxt = more ? load_in(k + 1) : In{};
    const real_t w[7] = {z[0], z[1], z[2], z[3], z[4], z[5], z[6]};
    cell(k, w, cur);
    cur = nxt;
#pragma unroll
    for (int s = 0; s < 7; s++) z[s] = z[s + 1];
  }
}

// FP32 stages that need both one-sided derivatives (normal motion and
// reinitialization).  The plus derivative at a cell and the minus derivative
// at the next cell along an axis share one WENO5 window (stencil::fastf::win5),
// so each window is formed once: along z the thread carries the window of
// plane k+1 to the next plane; along x and y the windows of the tile go
// through LDS, and the windows one cell past the tile are formed by a single
// wave (the last tile row for y, eight lanes of the first wave for x), so that
// no wave runs a divergent branch for its edge lane.
using stencil::fastf::Win;

// F is real_t (float): a template parameter so that the FP64 build, which never
// launches this kernel, does not compile its FP32-only body.
template <int RHS, int STAGE, class F>
__global__ void __launch_bounds__(TX* TY) fused_pm_f32_kernel(FusedArgs a) {
  using namespace stencil::fastf;
  __shared__ F plane[PLANE];
  // the plus derivative of the cell before (x) and the row below (y), formed by
  // the cell that owns the window it comes from; column TX of xplus: the
  // cells one past the tile
  __shared__ F yplus[TY][TX];
  __shared__ F xplus[TY][TX + 1];
  const int tid = threadIdx.y * TX + threadIdx.x;
  const int nt = a.nti * a.ntj * a.ntk;
  const int b = blockIdx.x;
  const int per = (nt + 7) / 8;
  const int t = (b % 8) * per + b / 8;
  if (t >= nt) return;
  const int ti = t % a.nti, tj = (t / a.nti) % a.ntj, tk = t / (a.nti * a.ntj);
  const int i0 = a.fb.ilo + ti * TX, j0 = a.fb.jlo + tj * TY;
  const int kb = a.fb.klo + tk * a.kc;
  const int ke = min(kb + a.kc - 1, a.fb.khi);
  const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
  const Box gb = a.gb;
  const bool active = i <= a.fb.ihi && j <= a.fb.jhi;
  const bool own_in = i <= gb.ihi && j <= gb.jhi;
  const long long own = (long long)(i - gb.ilo) + (long long)a.sy * (j - gb.jlo);
  long long hoff = -1;
  int hslot = 0;
  if (tid < NHALO) {
    int px, py;
    halo_pos(tid, px, py);
    const int gi = i0 - H + px, gj = j0 - H + py;
    hslot = py * PX + px;
    if (gi >= gb.ilo && gi <= gb.ihi && gj >= gb.jlo && gj <= gb.jhi)
      hoff = (long long)(gi - gb.ilo) + (long long)a.sy * (gj - gb.jlo);
  }
  const int lx = threadIdx.x + H, ly = threadIdx.y + H;
  const F* us = a.us;
  const long long sz = a.sz;
  const F ix = a.ix, iy = a.iy, iz = a.iz;

  F zw[7];  // zw[s+3] = u(i, j, k+s)
#pragma unroll
  for (int s = 0; s < 6; s++) zw[s] = own_in ? us[own + sz * (kb - H + s - gb.klo)] : F(0);
  F dz[5];  // dz[t] = D(k-1+t), D(m) = u(m) - u(m-1) along z
#pragma unroll
  for (int q = 0; q < 4; q++) dz[q] = zw[q + 2] - zw[q + 1];
  // minus derivative of the first plane, from its window D(k-2..k+2)
  F zm_next, zp_unused;
  fwd_rev(win5(zw[1] - zw[0], dz[0], dz[1], dz[2], dz[3]), dz[0], dz[1], dz[2], zm_next, zp_unused);
  zm_next *= iz;

  for (int k = kb; k <= ke; k++) {
    const long long zoff = sz * (k - gb.klo);
    zw[6] = own_in ? us[own + zoff + H * sz] : F(0);
    const F hv = hoff >= 0 ? us[hoff + zoff] : F(0);
    const long long off = own + zoff;
    const F ucv = STAGE != 1 && active ? a.uc[off] : F(0);
    F c0 = F(0);
    if (active) {
      if (RHS == kReinit) c0 = a.use_phi0 ? a.phi0[off] : F(0);
      else if (RHS == kNormalField) c0 = a.vnf[off];
    }
    __syncthreads();
    plane[ly * PX + lx] = zw[3];
    if (tid < NHALO) plane[hslot] = hv;
    __syncthreads();
    // z: the window D(k-1..k+3) gives plus(k) and minus(k+1); minus(k) was
    // formed from the previous plane's window
    dz[4] = zw[6] - zw[5];
    const Win wz1 = win5(dz[0], dz[1], dz[2], dz[3], dz[4]);
    float zpf, zmf;
    fwd_rev(wz1, dz[1], dz[2], dz[3], zmf, zpf);
    const float zp = zpf * iz;
    const float zm = zm_next;
    zm_next = zmf * iz;
    // x: this cell's window D(i-2..i+2) gives its minus derivative and the plus
    // derivative of the cell before it
    F dx_[6], dy_[6];
#pragma unroll
    for (int q = -2; q <= 3; q++) {
      dx_[q + 2] = plane[ly * PX + lx + q] - plane[ly * PX + lx + q - 1];
      dy_[q + 2] = plane[(ly + q) * PX + lx] - plane[(ly + q - 1) * PX + lx];
    }
    const Win wx = win5(dx_[0], dx_[1], dx_[2], dx_[3], dx_[4]);
    float xmf, xpf;
    fwd_rev(wx, dx_[1], dx_[2], dx_[3], xmf, xpf);
    const float xm = xmf * ix;
    xplus[threadIdx.y][threadIdx.x] = xpf * ix;
    if (threadIdx.y == 0 && threadIdx.x < TY) {  // one wave forms the TY plus derivatives of the tile's last column
      const int row = (threadIdx.x + H) * PX + TX + H;
      F e[5];
#pragma unroll
      for (int q = -2; q <= 2; q++) e[q + 2] = plane[row + q] - plane[row + q - 1];
      float ef, er;
      fwd_rev(win5(e[0], e[1], e[2], e[3], e[4]), e[1], e[2], e[3], ef, er);
      xplus[threadIdx.x][TX] = er * ix;
    }
    // y: likewise along j
    const Win wy = win5(dy_[0], dy_[1], dy_[2], dy_[3], dy_[4]);
    float ymf, ypf;
    fwd_rev(wy, dy_[1], dy_[2], dy_[3], ymf, ypf);
    const float ym = ymf * iy;
    yplus[threadIdx.y][threadIdx.x] = ypf * iy;
    __syncthreads();
    float yp;
    if (threadIdx.y == TY - 1) {  // the last row forms the plus derivative from the row past the tile
      float nf, nr;
      fwd_rev(win5(dy_[1], dy_[2], dy_[3], dy_[4], dy_[5]), dy_[2], dy_[3], dy_[4], nf, nr);
      yp = nr * iy;
    } else {
      yp = yplus[threadIdx.y + 1][threadIdx.x];
    }
    const float xp = xplus[threadIdx.y][threadIdx.x + 1];
    if (active) {
      F r;
      if (RHS == kReinit) {
        const F s0 = a.use_phi0 ? c0 : zw[3];
        r = reinit_rhs(s0, a.use_phi0 != 0, xp, yp, zp, xm, ym, zm, a.dx_sq);
      } else {
        const F v = RHS == kNormalField ? c0 : a.vn;
        r = 0.f;
        if (fabsf(v) >= kZeroTol)
          r = r - v * __builtin_amdgcn_sqrtf(stencil::godunov_sq(v > 0.f, xp, yp, zp, xm, ym, zm));
      }
      const F uc = STAGE == 1 ? zw[3] : ucv;
      a.out[off] = rk_combine<STAGE>(uc, zw[3], a.dt, r);
    }
#pragma unroll
    for (int s = 0; s < 6; s++) zw[s] = zw[s + 1];
#pragma unroll
    for (int q = 0; q < 4; q++) dz[q] = dz[q + 1];
  }
}

// ---------------------------------------------------------------------------
// FP32 advection stages with the WENO5 inputs shared between neighbouring
// cells.
//
// Along an axis, write D(m) = u(m) - u(m-1) and, for the three consecutive
// differences around position m, A(m) = D(m-1) - 2 D(m) + D(m+1),
// C(m) = D(m-1) - D(m+1) and T(m) = 13/3 A(m)^2.  The minus-side
// reconstruction at m, weno5(D(m-2), .., D(m+2)), needs (4 S_k = 13/3 a_k^2 +
// b_k^2, the scaling of stencil::fastf::win5)
//   4 S1 = T(m-1) + (2 A(m-1) - C(m-1))^2,  4 S2 = T(m) + C(m)^2,
//   4 S3 = T(m+1) + (2 A(m+1) + C(m+1))^2,
//   a1 - a2 = A(m-1) - A(m),  a2 - a3 = A(m) - A(m+1),
//   p2 = D(m) + A(m)/12 - C(m)/4,
// and the plus-side one, weno5(D(m+3), .., D(m-1)), the same terms at
// m+2, m+1, m with the sign of C reversed.  So each position's D, A, C, T are
// formed once and read by every cell whose stencil holds them: along z from
// registers carried from plane to plane, along x and y from LDS planes (x: each
// wave forms its own row, the five positions past the tile by a few lanes;
// y: the tile's rows and the rows past it, shared out over the waves).  The
// upwind side is a branch rather than a per-lane select: waves whose cells all
// see the same sign along an axis (nearly all of them for a smooth velocity)
// evaluate one side only.  Same weights and candidates as the other FP32
// forms (stencil::fastf), different rounding.
// ---------------------------------------------------------------------------
constexpr int AXD = TX + 5;  // D along x at i0-2 .. i0+TX+2
constexpr int AXQ = TX + 3;  // A, C, T along x at i0-1 .. i0+TX+1
constexpr int AYD = TY + 5;  // D along y at rows j0-2 .. j0+TY+2
constexpr int AYQ = TY + 3;  // A, C, T along y at rows j0-1 .. j0+TY+1

struct Quad {  // A, C, T of one position
  float A, C, T;
};
__device__ __forceinline__ Quad quad(float dm, float d0, float dp) {
  const float A = fmaf(-2.f, d0, dm) + dp;
  return Quad{A, dm - dp, (13.f / 3.f) * A * A};
}

// weno5 from 4 S_k (unnormalised), the candidate differences d12 = a1 - a2,
// d23 = a2 - a3, the middle candidate p2 and ma = max |v|: win5 + fwd of
// stencil::fastf with the terms given.  With Q_k = e_k^2, the weighted sum
//   p2 + (0.1 Q2Q3 d12/3 + 0.3 Q1Q2 d23/6) / (0.1 Q2Q3 + 0.6 Q1Q3 + 0.3 Q1Q2)
// is formed as p2 + Q2 (Q3 d12/3 + Q1 d23/2) / (Q2 (Q3 + 3 Q1) + 6 Q1 Q3).
__device__ __forceinline__ float weno_core(float E1, float E2, float E3, float d12, float d23, float p2, float ma) {
  using stencil::fastf::rcp;
  const float im = rcp(fmaf(ma, ma, 1e-30f));
  const float e1 = fmaf(E1, im, 4e-6f), e2 = fmaf(E2, im, 4e-6f), e3 = fmaf(E3, im, 4e-6f);
  const float Q1 = e1 * e1, Q2 = e2 * e2, Q3 = e3 * e3;
  const float num = Q2 * fmaf(Q3 * d12, 1.f / 3.f, 0.5f * (Q1 * d23));
  const float den = fmaf(6.f * Q1, Q3, Q2 * fmaf(3.f, Q1, Q3));
  return fmaf(num, rcp(den), p2);
}

// max |v| over five values as a chain (two v_max3 with absolute-value operands)
__device__ __forceinline__ float amax5(float a, float b, float c, float d, float e) {
  return fmaxf(fmaxf(fmaxf(fmaxf(fabsf(a), fabsf(b)), fabsf(c)), fabsf(d)), fabsf(e));
}

// The upwinded reconstruction (in undivided units) at position m of one axis
// from D(m-2..m+3) (d[0], d[DS], .., d[5 DS]) and the quads at m-1..m+2 (q[0],
// q[QS], .., q[3 QS]): the minus side for vel > 0, the plus side otherwise, 0
// when |vel| < tol.  Each side reads and finishes only its own terms, so a
// wave of one sign loads and evaluates one side.
template <int DS, int QS, class DP, class QP>
__device__ __forceinline__ float upwind_q(float vel, float tol, DP d, QP q) {
  float r = 0.f;
  if (!(fabsf(vel) < tol)) {
    if (vel > 0.f) {
      const Quad q0 = q[0], q1 = q[QS], q2 = q[2 * QS];
      const float b1 = fmaf(2.f, q0.A, -q0.C), b3 = fmaf(2.f, q2.A, q2.C);
      r = weno_core(fmaf(b1, b1, q0.T), fmaf(q1.C, q1.C, q1.T), fmaf(b3, b3, q2.T), q0.A - q1.A, q1.A - q2.A,
                    fmaf(-0.25f, q1.C, fmaf(1.f / 12.f, q1.A, d[2 * DS])),
                    amax5(d[0], d[DS], d[2 * DS], d[3 * DS], d[4 * DS]));
    } else {
      const Quad q1 = q[QS], q2 = q[2 * QS], q3 = q[3 * QS];
      const float b1 = fmaf(2.f, q3.A, q3.C), b3 = fmaf(2.f, q1.A, -q1.C);
      r = weno_core(fmaf(b1, b1, q3.T), fmaf(q2.C, q2.C, q2.T), fmaf(b3, b3, q1.T), q3.A - q2.A, q2.A - q1.A,
                    fmaf(0.25f, q2.C, fmaf(1.f / 12.f, q2.A, d[3 * DS])),
                    amax5(d[DS], d[2 * DS], d[3 * DS], d[4 * DS], d[5 * DS]));
    }
  }
  return r;
}

template <int RHS, int STAGE, class F>
__global__ void __launch_bounds__(TX* TY) fused_advect_f32_kernel(FusedArgs a) {
  __shared__ F plane[PLANE];       // u at plane k with a 3-cell halo, [py][px]
  __shared__ F dxs[TY][AXD];       // D along x, index i - i0 + 2
  __shared__ Quad qxs[TY][AXQ];    // quads along x, index i - i0 + 1
  __shared__ F dys[AYD][TX];       // D along y, row j - j0 + 2
  __shared__ Quad qys[AYQ][TX];    // quads along y, row j - j0 + 1
  static_assert(AYD - TY <= TY && AYQ - TY <= TY && TY * (AXD - TX) <= TX && TY * (AXQ - TX) <= TX && TY >= 7,
                "one pass of waves covers the rows past the tile; waves 5 and 6 the positions past each row");
  const int tid = threadIdx.y * TX + threadIdx.x;
  const int nt = a.nti * a.ntj * a.ntk;
  const int b = blockIdx.x;
  const int per = (nt + 7) / 8;
  const int t = (b % 8) * per + b / 8;
  if (t >= nt) return;
  const int ti = t % a.nti, tj = (t / a.nti) % a.ntj, tk = t / (a.nti * a.ntj);
  const int i0 = a.fb.ilo + ti * TX, j0 = a.fb.jlo + tj * TY;
  const int kb = a.fb.klo + tk * a.kc;
  const int ke = min(kb + a.kc - 1, a.fb.khi);
  const int lx = threadIdx.x, ly = threadIdx.y;
  const int i = i0 + lx, j = j0 + ly;
  const Box gb = a.gb;
  const bool active = i <= a.fb.ihi && j <= a.fb.jhi;
  const bool own_in = i <= gb.ihi && j <= gb.jhi;
  const long long own = (long long)(i - gb.ilo) + (long long)a.sy * (j - gb.jlo);
  long long hoff = -1;
  int hslot = 0;
  if (tid < NHALO) {
    int px, py;
    halo_pos(tid, px, py);
    const int gi = i0 - H + px, gj = j0 - H + py;
    hslot = py * PX + px;
    if (gi >= gb.ilo && gi <= gb.ihi && gj >= gb.jlo && gj <= gb.jhi)
      hoff = (long long)(gi - gb.ilo) + (long long)a.sy * (gj - gb.jlo);
  }
  const F* us = a.us;
  const long long sz = a.sz;
  // positions past each row along x: wave 5 forms D at dxs index 0, 1 (i0-2,
  // i0-1) and TX+2 .. TX+4 of every row (lane = 5 row + slot), wave 6 the
  // quads at qxs index 0, TX+1, TX+2 (lane = 3 row + slot)
  const int hd_row = lx / (AXD - TX), hd_slot = lx % (AXD - TX);
  const int hd = hd_slot < 2 ? hd_slot : TX + hd_slot;
  const int hq_row = lx / (AXQ - TX), hq_slot = lx % (AXQ - TX);
  const int hq = hq_slot == 0 ? 0 : TX + hq_slot;
  const bool hd_lane = ly == 5 && lx < TY * (AXD - TX), hq_lane = ly == 6 && lx < TY * (AXQ - TX);

  // z: u at k..k+2, D at k-2..k+3 (dz[s+2] = D(k+s)), quads at k-1..k+2 (qz[s+1])
  F u0, u1, u2;
  F dz[6];
  Quad qz[4];
  {
    F w[6];
#pragma unroll
    for (int s = 0; s < 6; s++) w[s] = own_in ? us[own + sz * (kb - H + s - gb.klo)] : F(0);  // k-3 .. k+2
#pragma unroll
    for (int s = 0; s < 5; s++) dz[s] = w[s + 1] - w[s];  // D(k-2 .. k+2)
#pragma unroll
    for (int s = 0; s < 3; s++) qz[s] = quad(dz[s], dz[s + 1], dz[s + 2]);  // k-1 .. k+1
    u0 = w[3];
    u1 = w[4];
    u2 = w[5];
  }
  F ta0 = 0, tb0 = 0, ta1 = 0, tb1 = 0;
  if (RHS == kAdvectDeform && active) {
    ta0 = a.ta[0][i - gb.ilo];
    tb0 = a.tb[0][i - gb.ilo];
    ta1 = a.ta[1][j - gb.jlo];
    tb1 = a.tb[1][j - gb.jlo];
  }
  // the deformation velocity divided by the grid spacing: u/dx = X0 B(z), v/dy = X1 B(z), w/dz = X2 A(z)
  const F X0 = F(2) * a.tfac * a.ix * ta0 * tb1, X1 = -(a.tfac * a.iy) * tb0 * ta1, X2 = -(a.tfac * a.iz) * tb0 * tb1;
  const F tolx = kZeroTol * a.ix, toly = kZeroTol * a.iy, tolz = kZeroTol * a.iz;

  for (int k = kb; k <= ke; k++) {
    const long long zoff = sz * (k - gb.klo);
    const F u3 = own_in ? us[own + zoff + H * sz] : F(0);
    const F hv = hoff >= 0 ? us[hoff + zoff] : F(0);
    const long long off = own + zoff;
    const F ucv = STAGE != 1 && active ? a.uc[off] : F(0);
    F c0 = F(0), c1 = F(0), c2 = F(0);
    if (RHS == kAdvectField && active) {
      c0 = a.vx[off];
      c1 = a.vy[off];
      c2 = a.vz[off];
    }
    __syncthreads();  // the previous plane's LDS reads are done
    plane[(ly + H) * PX + lx + H] = u0;
    if (tid < NHALO) plane[hslot] = hv;
    // z: D(k+3) and the quad at k+2
    dz[5] = u3 - u2;
    qz[3] = quad(dz[3], dz[4], dz[5]);
    __syncthreads();
    // D along x (own row) and along y (row ly, and the rows past the tile)
    dxs[ly][lx + 2] = plane[(ly + H) * PX + lx + H] - plane[(ly + H) * PX + lx + H - 1];
    if (hd_lane) dxs[hd_row][hd] = plane[(hd_row + H) * PX + hd + 1] - plane[(hd_row + H) * PX + hd];
    dys[ly][lx] = plane[(ly + 1) * PX + lx + H] - plane[ly * PX + lx + H];
    if (ly < AYD - TY) dys[ly + TY][lx] = plane[(ly + TY + 1) * PX + lx + H] - plane[(ly + TY) * PX + lx + H];
    __syncthreads();
    qxs[ly][lx + 1] = quad(dxs[ly][lx + 1], dxs[ly][lx + 2], dxs[ly][lx + 3]);
    if (hq_lane) qxs[hq_row][hq] = quad(dxs[hq_row][hq], dxs[hq_row][hq + 1], dxs[hq_row][hq + 2]);
    qys[ly][lx] = quad(dys[ly][lx], dys[ly + 1][lx], dys[ly + 2][lx]);
    if (ly < AYQ - TY) qys[ly + TY][lx] = quad(dys[ly + TY][lx], dys[ly + TY + 1][lx], dys[ly + TY + 2][lx]);
    __syncthreads();
    if (active) {
      F ux, vy, wz;
      if (RHS == kAdvectDeform) {
        ux = X0 * a.tb[2][k - gb.klo];
        vy = X1 * a.tb[2][k - gb.klo];
        wz = X2 * a.ta[2][k - gb.klo];
      } else {
        ux = c0 * a.ix;
        vy = c1 * a.iy;
        wz = c2 * a.iz;
      }
      const F px = upwind_q<1, 1>(ux, tolx, &dxs[ly][lx], &qxs[ly][lx]);
      const F py = upwind_q<TX, TX>(vy, toly, &dys[ly][lx], &qys[ly][lx]);
      const F pz = upwind_q<1, 1>(wz, tolz, dz, qz);
      // r = -(u phi_x + v phi_y + w phi_z), the velocities already divided by the spacing
      const F r = F(0) - (ux * px + vy * py + wz * pz);
      const F uc = STAGE == 1 ? u0 : ucv;
      a.out[off] = rk_combine<STAGE>(uc, u0, a.dt, r);
    }
    u0 = u1;
    u1 = u2;
    u2 = u3;
#pragma unroll
    for (int s = 0; s < 5; s++) dz[s] = dz[s + 1];
#pragma unroll
    for (int s = 0; s < 3; s++) qz[s] = qz[s + 1];
  }
}

// ---------------------------------------------------------------------------
// FP32 stages that need both one-sided derivatives (normal motion and
// reinitialization), on the shared quads of fused_advect_f32_kernel.  The
// minus derivative at m and the plus derivative at m-1 are the two orders of
// one window D(m-2..m+2): its smoothness terms (from the quads at m-1, m, m+1)
// are formed once and each order finishes from them.  Along x and y each cell
// forms its own window, keeps the minus derivative and hands the plus
// derivative of the cell before it through LDS (the windows one cell past
// the tile are formed by waves 6 and 7); along z the window of plane k+1
// gives this plane's plus derivative and, carried, the next plane's minus.
// ---------------------------------------------------------------------------
struct Window {  // the part of window W(m) both orders read
  float Q1, Q2, Q3, Q3d12, Q1d23, S6, t, C;
};
// d: D(m-2..m+2); quads at m-1, m, m+1
__device__ __forceinline__ Window window(float d0, float d1, float d2, float d3, float d4, const Quad& qm, const Quad& q0,
                                      const Quad& qp) {
  using stencil::fastf::rcp;
  const float b1 = fmaf(2.f, qm.A, -qm.C), b3 = fmaf(2.f, qp.A, qp.C);
  const float im = rcp(fmaf(amax5(d0, d1, d2, d3, d4), amax5(d0, d1, d2, d3, d4), 1e-30f));
  const float e1 = fmaf(fmaf(b1, b1, qm.T), im, 4e-6f), e2 = fmaf(fmaf(q0.C, q0.C, q0.T), im, 4e-6f),
              e3 = fmaf(fmaf(b3, b3, qp.T), im, 4e-6f);
  Window w;
  w.Q1 = e1 * e1;
  w.Q2 = e2 * e2;
  w.Q3 = e3 * e3;
  w.Q3d12 = w.Q3 * (qm.A - q0.A);
  w.Q1d23 = w.Q1 * (q0.A - qp.A);
  w.S6 = (6.f * w.Q1) * w.Q3;
  w.t = fmaf(1.f / 12.f, q0.A, d2);
  w.C = q0.C;
  return w;
}
// minus derivative at m (undivided): weno5(D(m-2), .., D(m+2))
__device__ __forceinline__ float win_minus(const Window& w) {
  const float num = w.Q2 * fmaf(w.Q3d12, 1.f / 3.f, 0.5f * w.Q1d23);
  const float den = fmaf(w.Q2, fmaf(3.f, w.Q1, w.Q3), w.S6);
  return fmaf(num, stencil::fastf::rcp(den), fmaf(-0.25f, w.C, w.t));
}
// plus derivative at m-1 (undivided): weno5(D(m+2), .., D(m-2))
__device__ __forceinline__ float win_plus(const Window& w) {
  const float num = -(w.Q2 * fmaf(w.Q1d23, 1.f / 3.f, 0.5f * w.Q3d12));
  const float den = fmaf(w.Q2, fmaf(3.f, w.Q3, w.Q1), w.S6);
  return fmaf(num, stencil::fastf::rcp(den), fmaf(0.25f, w.C, w.t));
}

// Godunov upwinding of one axis from the undivided one-sided derivatives:
// max(max(m, 0)^2, min(p, 0)^2) = max(m, -p, 0)^2 for a positive speed (or
// for reinitialization where the sign source is positive), max(-m, p, 0)^2
// otherwise; the 1/dx of the derivatives applied to the selected value.
__device__ __forceinline__ float godunov_axis(bool pos, float p, float m, float inv) {
  const float s = pos ? fmaxf(fmaxf(m, -p), 0.f) : fmaxf(fmaxf(-m, p), 0.f);
  return s * inv;
}

template <int RHS, int STAGE, class F>
__global__ void __launch_bounds__(TX* TY) fused_pm2_f32_kernel(FusedArgs a) {
  __shared__ F plane[PLANE];
  __shared__ F dxs[TY][AXD];
  __shared__ Quad qxs[TY][AXQ];
  __shared__ F dys[AYD][TX];
  __shared__ Quad qys[AYQ][TX];
  __shared__ F xplus[TY][TX];  // plus derivative along x (undivided), handed from the cell after
  __shared__ F yplus[TY][TX];
  static_assert(AYD - TY <= TY && AYQ - TY <= TY && TY * (AXD - TX) <= TX && TY * (AXQ - TX) <= TX && TY >= 8,
                "one pass of waves covers the rows past the tile; waves 5 .. 7 the positions past each row");
  const int tid = threadIdx.y * TX + threadIdx.x;
  const int nt = a.nti * a.ntj * a.ntk;
  const int b = blockIdx.x;
  const int per = (nt + 7) / 8;
  const int t = (b % 8) * per + b / 8;
  if (t >= nt) return;
  const int ti = t % a.nti, tj = (t / a.nti) % a.ntj, tk = t / (a.nti * a.ntj);
  const int i0 = a.fb.ilo + ti * TX, j0 = a.fb.jlo + tj * TY;
  const int kb = a.fb.klo + tk * a.kc;
  const int ke = min(kb + a.kc - 1, a.fb.khi);
  const int lx = threadIdx.x, ly = threadIdx.y;
  const int i = i0 + lx, j = j0 + ly;
  const Box gb = a.gb;
  const bool active = i <= a.fb.ihi && j <= a.fb.jhi;
  const bool own_in = i <= gb.ihi && j <= gb.jhi;
  const long long own = (long long)(i - gb.ilo) + (long long)a.sy * (j - gb.jlo);
  long long hoff = -1;
  int hslot = 0;
  if (tid < NHALO) {
    int px, py;
    halo_pos(tid, px, py);
    const int gi = i0 - H + px, gj = j0 - H + py;
    hslot = py * PX + px;
    if (gi >= gb.ilo && gi <= gb.ihi && gj >= gb.jlo && gj <= gb.jhi)
      hoff = (long long)(gi - gb.ilo) + (long long)a.sy * (gj - gb.jlo);
  }
  const F* us = a.us;
  const long long sz = a.sz;
  const F ix = a.ix, iy = a.iy, iz = a.iz;
  const int hd_row = lx / (AXD - TX), hd_slot = lx % (AXD - TX);
  const int hd = hd_slot < 2 ? hd_slot : TX + hd_slot;
  const int hq_row = lx / (AXQ - TX), hq_slot = lx % (AXQ - TX);
  const int hq = hq_slot == 0 ? 0 : TX + hq_slot;
  const bool hd_lane = ly == 5 && lx < TY * (AXD - TX), hq_lane = ly == 6 && lx < TY * (AXQ - TX);

  F u0, u1, u2;
  F dz[6];
  Quad qz[4];
  F zminus;  // minus derivative along z at plane k (undivided)
  {
    F w[6];
#pragma unroll
    for (int s = 0; s < 6; s++) w[s] = own_in ? us[own + sz * (kb - H + s - gb.klo)] : F(0);  // k-3 .. k+2
#pragma unroll
    for (int s = 0; s < 5; s++) dz[s] = w[s + 1] - w[s];  // D(k-2 .. k+2)
#pragma unroll
    for (int s = 0; s < 3; s++) qz[s] = quad(dz[s], dz[s + 1], dz[s + 2]);  // k-1 .. k+1
    u0 = w[3];
    u1 = w[4];
    u2 = w[5];
    zminus = win_minus(window(dz[0], dz[1], dz[2], dz[3], dz[4], qz[0], qz[1], qz[2]));
  }

  for (int k = kb; k <= ke; k++) {
    const long long zoff = sz * (k - gb.klo);
    const F u3 = own_in ? us[own + zoff + H * sz] : F(0);
    const F hv = hoff >= 0 ? us[hoff + zoff] : F(0);
    const long long off = own + zoff;
    const F ucv = STAGE != 1 && active ? a.uc[off] : F(0);
    F c0 = F(0);
    if (active) {
      if (RHS == kReinit) c0 = a.use_phi0 ? a.phi0[off] : F(0);
      else if (RHS == kNormalField) c0 = a.vnf[off];
    }
    __syncthreads();  // the previous plane's LDS reads are done
    plane[(ly + H) * PX + lx + H] = u0;
    if (tid < NHALO) plane[hslot] = hv;
    dz[5] = u3 - u2;
    qz[3] = quad(dz[3], dz[4], dz[5]);
    // z: the window of plane k+1 -- this plane's plus derivative and the next one's minus
    const Window wz = window(dz[1], dz[2], dz[3], dz[4], dz[5], qz[1], qz[2], qz[3]);
    const F zp = win_plus(wz) * iz, zm = zminus * iz;
    zminus = win_minus(wz);
    __syncthreads();
    dxs[ly][lx + 2] = plane[(ly + H) * PX + lx + H] - plane[(ly + H) * PX + lx + H - 1];
    if (hd_lane) dxs[hd_row][hd] = plane[(hd_row + H) * PX + hd + 1] - plane[(hd_row + H) * PX + hd];
    dys[ly][lx] = plane[(ly + 1) * PX + lx + H] - plane[ly * PX + lx + H];
    if (ly < AYD - TY) dys[ly + TY][lx] = plane[(ly + TY + 1) * PX + lx + H] - plane[(ly + TY) * PX + lx + H];
    __syncthreads();
    qxs[ly][lx + 1] = quad(dxs[ly][lx + 1], dxs[ly][lx + 2], dxs[ly][lx + 3]);
    if (hq_lane) qxs[hq_row][hq] = quad(dxs[hq_row][hq], dxs[hq_row][hq + 1], dxs[hq_row][hq + 2]);
    qys[ly][lx] = quad(dys[ly][lx], dys[ly + 1][lx], dys[ly + 2][lx]);
    if (ly < AYQ - TY) qys[ly + TY][lx] = quad(dys[ly + TY][lx], dys[ly + TY + 1][lx], dys[ly + TY + 2][lx]);
    __syncthreads();
    // x and y: this cell's windows; the plus derivative of the cell before goes through LDS
    const F* dr = dxs[ly];
    const Window wx = window(dr[lx], dr[lx + 1], dr[lx + 2], dr[lx + 3], dr[lx + 4], qxs[ly][lx], qxs[ly][lx + 1],
                          qxs[ly][lx + 2]);
    const Window wy = window(dys[ly][lx], dys[ly + 1][lx], dys[ly + 2][lx], dys[ly + 3][lx], dys[ly + 4][lx], qys[ly][lx],
                          qys[ly + 1][lx], qys[ly + 2][lx]);
    if (lx > 0) xplus[ly][lx - 1] = win_plus(wx);
    if (ly > 0) yplus[ly - 1][lx] = win_plus(wy);
    if (ly == 6 && lx < TY) {  // the window one cell past each row along x
      const F* e = dxs[lx];
      xplus[lx][TX - 1] = win_plus(window(e[TX], e[TX + 1], e[TX + 2], e[TX + 3], e[TX + 4], qxs[lx][TX],
                                          qxs[lx][TX + 1], qxs[lx][TX + 2]));
    }
    if (ly == 7)  // the window one row past the tile along y
      yplus[TY - 1][lx] = win_plus(window(dys[TY][lx], dys[TY + 1][lx], dys[TY + 2][lx], dys[TY + 3][lx],
                                          dys[TY + 4][lx], qys[TY][lx], qys[TY + 1][lx], qys[TY + 2][lx]));
    const F xm = win_minus(wx) * ix, ym = win_minus(wy) * iy;
    __syncthreads();
    if (active) {
      const F xp = xplus[ly][lx] * ix, yp = yplus[ly][lx] * iy;
      F r;
      if (RHS == kReinit) {
        const F s = a.use_phi0 ? c0 : u0;
        const bool pos = s > F(0);
        const F gx = godunov_axis(pos, xp, xm, F(1)), gy = godunov_axis(pos, yp, ym, F(1)),
                gz = godunov_axis(pos, zp, zm, F(1));
        const F n2 = fmaf(gz, gz, fmaf(gy, gy, gx * gx));
        r = F(0);
        if (fabsf(s) >= kZeroTol) {
          const F sgn = s * __builtin_amdgcn_rsqf(fmaf(s, s, a.use_phi0 ? a.dx_sq : n2 * a.dx_sq));
          r = sgn * (F(1) - __builtin_amdgcn_sqrtf(n2));
        }
      } else {
        const F v = RHS == kNormalField ? c0 : a.vn;
        const bool pos = v > F(0);
        const F gx = godunov_axis(pos, xp, xm, F(1)), gy = godunov_axis(pos, yp, ym, F(1)),
                gz = godunov_axis(pos, zp, zm, F(1));
        r = F(0);
        if (fabsf(v) >= kZeroTol) r = r - v * __builtin_amdgcn_sqrtf(fmaf(gz, gz, fmaf(gy, gy, gx * gx)));
      }
      const F uc = STAGE == 1 ? u0 : ucv;
      a.out[off] = rk_combine<STAGE>(uc, u0, a.dt, r);
    }
    u0 = u1;
    u1 = u2;
    u2 = u3;
#pragma unroll
    for (int s = 0; s < 5; s++) dz[s] = dz[s + 1];
#pragma unroll
    for (int s = 0; s < 3; s++) qz[s] = qz[s + 1];
  }
}

// FP32 stage kernels: 1 (default) = the per-cell window kernels
// (fused_stage_kernel, fused_pm_f32_kernel), 2 = the shared-quantity kernels
// (fused_advect_f32_kernel, fused_pm2_f32_kernel).  The second set issues
// 14-20 % fewer VALU instructions per cell but holds 72-80 VGPRs and meets
// four barriers per plane; at the 6 waves per SIMD that allows it runs slower
// on gfx950 (0.91 vs 0.76 ms per 512^3 advection stage).  It stays as an
// independent rounding order for precision comparisons.  FP64 always runs
// fused_stage_kernel.
int g_f32_kernels = 1;

template <int RHS, int STAGE>
void launch_stage(const FusedArgs& a, int nblocks) {
  if constexpr (sizeof(real_t) == 4 && (RHS == kNormalConst || RHS == kNormalField || RHS == kReinit)) {
    if (g_f32_kernels == 2)
      hipLaunchKernelGGL((fused_pm2_f32_kernel<RHS, STAGE, real_t>), dim3(nblocks), dim3(TX, TY), 0, stream(), a);
    else
      hipLaunchKernelGGL((fused_pm_f32_kernel<RHS, STAGE, real_t>), dim3(nblocks), dim3(TX, TY), 0, stream(), a);
  } else if constexpr (sizeof(real_t) == 4 && (RHS == kAdvectField || RHS == kAdvectDeform)) {
    if (g_f32_kernels == 2)
      hipLaunchKernelGGL((fused_advect_f32_kernel<RHS, STAGE, real_t>), dim3(nblocks), dim3(TX, TY), 0, stream(), a);
    else
      hipLaunchKernelGGL((fused_stage_kernel<RHS, STAGE>), dim3(nblocks), dim3(TX, TY), 0, stream(), a);
  } else
    hipLaunchKernelGGL((fused_stage_kernel<RHS, STAGE>), dim3(nblocks), dim3(TX, TY), 0, stream(), a);
}

template <int RHS>
bool launch_rhs(int stage, const FusedArgs& a, int nblocks) {
  switch (stage) {
    case 1: launch_stage<RHS, 1>(a, nblocks); break;
    case 2: launch_stage<RHS, 2>(a, nblocks); break;
    case 3: launch_stage<RHS, 3>(a, nblocks); break;
    case 22: launch_stage<RHS, 22>(a, nblocks); break;
    default: return false;
  }
  return true;
}

int g_kchunk = 64;

int run(const char* who, FusedArgs a, int stage) {
  if (a.fb.empty()) return 0;
  if (!check_box(a.gb, a.fb.widen(H, H, H), who, "phi") || !check_size(a.gb, who)) return -1;
  if (stage == 1 && a.uc == nullptr) a.uc = a.us;
  if (stage != 1 && a.uc == nullptr) {
    set_error(kErrArg, "%s: stage %d needs phi at the start of the step", who, stage);
    return -1;
  }
  a.sy = a.gb.nx();
  a.sz = a.gb.nx() * a.gb.ny();
  a.kc = g_kchunk;
  a.nti = (a.fb.nx() + TX - 1) / TX;
  a.ntj = (a.fb.ny() + TY - 1) / TY;
  a.ntk = (a.fb.nz() + a.kc - 1) / a.kc;
  const long long nt = (long long)a.nti * a.ntj * a.ntk;
  const int per = (int)((nt + 7) / 8);
  const int nblocks = 8 * per;  // surplus blocks of the last XCD range exit at once
  bool ok = false;
  switch (a.rhs) {
    case kNormalConst: ok = launch_rhs<kNormalConst>(stage, a, nblocks); break;
    case kNormalField: ok = launch_rhs<kNormalField>(stage, a, nblocks); break;
    case kAdvectField: ok = launch_rhs<kAdvectField>(stage, a, nblocks); break;
    case kAdvectDeform: ok = launch_rhs<kAdvectDeform>(stage, a, nblocks); break;
    case kReinit: ok = launch_rhs<kReinit>(stage, a, nblocks); break;
  }
  if (!ok) {
    set_error(kErrArg, "%s: stage must be 1, 2, 3 (TVD-RK3) or 22 (TVD-RK2 stage 2), got %d", who, stage);
    return -1;
  }
  return finish(who) ? 0 : -1;
}

FusedArgs base(real_t* out, const real_t* us, const real_t* uc, const Box& gb, const Box& fb, real_t dt, real_t dx,
               real_t dy, real_t dz) {
  FusedArgs a{};
  a.us = us;
  a.uc = uc;
  a.out = out;
  a.gb = gb;
  a.fb = fb;
  a.dt = dt;
  a.ix = real_t(1.0) / dx;
  a.iy = real_t(1.0) / dy;
  a.iz = real_t(1.0) / dz;
  return a;
}

__global__ void __launch_bounds__(256) deform_fill_kernel(Box b, Box gb, real_t* vx, real_t* vy, real_t* vz,
                                                          const real_t* ta0, const real_t* ta1, const real_t* ta2,
                                                          const real_t* tb0, const real_t* tb1, const real_t* tb2,
                                                          real_t c) {
  LSMHIP_CELL(b);
  const real_t* ta[3] = {ta0, ta1, ta2};
  const real_t* tb[3] = {tb0, tb1, tb2};
  real_t u, v, w;
  deform_velocity(ta, tb, i - gb.ilo, j - gb.jlo, k - gb.klo, c, u, v, w);
  const int off = (i - gb.ilo) + gb.nx() * ((j - gb.jlo) + gb.ny() * (k - gb.klo));
  vx[off] = u;
  vy[off] = v;
  vz[off] = w;
}

}  // namespace

// ---------------------------------------------------------------------------
// Fused stage entry points (no counterpart in the reference; the per-cell
// arithmetic is that of the drop-in chain named in each comment).  All arrays
// are device arrays on the ghost box gb; the stage writes phi_out on fb only.
// stage: 1, 2, 3 = TVD-RK3 stages (lsm3dTVDRK3Stage1/2/3), 22 = TVD-RK2 stage 2.
// phi_n may be null for stage 1.  Returns 0, or -1 with lsmhip_get_last_error set.
// ---------------------------------------------------------------------------

// HJ_WENO5 -> ADD_[CONST_]NORMAL_VEL_TERM -> RK stage.  vn_field null: constant speed vn.
LSMHIP_EXPORT int lsmhip_weno5_stage_normal(real_t* phi_out, const real_t* phi_stage, const real_t* phi_n,
                                            LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), int stage, real_t vn,
                                            const real_t* vn_field, real_t dt, real_t dx, real_t dy, real_t dz) {
  LSMHIP_REQUIRE_ARRAYS_OR(-1, phi_out, phi_stage);
  FusedArgs a = base(phi_out, phi_stage, phi_n, LSMHIP_BOX(gb), LSMHIP_BOX(fb), dt, dx, dy, dz);
  a.rhs = vn_field ? kNormalField : kNormalConst;
  a.vn = vn;
  a.vnf = vn_field;
  if (!vn_field && !(fabs(vn) >= kZeroTol)) a.vn = real_t(0);
  return run("lsmhip_weno5_stage_normal", a, stage);
}

// UPWIND_HJ_WENO5 -> ADD_ADVECTION_TERM -> RK stage, velocity arrays on gb.
LSMHIP_EXPORT int lsmhip_weno5_stage_advect(real_t* phi_out, const real_t* phi_stage, const real_t* phi_n,
                                            LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), int stage, const real_t* vx,
                                            const real_t* vy, const real_t* vz, real_t dt, real_t dx, real_t dy,
                                            real_t dz) {
  LSMHIP_REQUIRE_ARRAYS_OR(-1, phi_out, phi_stage, vx, vy, vz);
  FusedArgs a = base(phi_out, phi_stage, phi_n, LSMHIP_BOX(gb), LSMHIP_BOX(fb), dt, dx, dy, dz);
  a.rhs = kAdvectField;
  a.vx = vx;
  a.vy = vy;
  a.vz = vz;
  return run("lsmhip_weno5_stage_advect", a, stage);
}

// As lsmhip_weno5_stage_advect with the divergence-free deformation velocity
//   u = 2 sin^2(pi x) sin(2 pi y) sin(2 pi z) c,  v = -sin(2 pi x) sin^2(pi y) sin(2 pi z) c,
//   w = -sin(2 pi x) sin(2 pi y) sin^2(pi z) c
// evaluated in the kernel from per-axis tables (device arrays indexed from the
// ghost-box origin: tabs = {A_x, A_y, A_z, B_x, B_y, B_z}, A = sin^2(pi x),
// B = sin(2 pi x)); c is the time factor.  No velocity array is read.
LSMHIP_EXPORT int lsmhip_weno5_stage_advect_deform(real_t* phi_out, const real_t* phi_stage, const real_t* phi_n,
                                                   LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), int stage,
                                                   const real_t* const* tabs, real_t c, real_t dt, real_t dx,
                                                   real_t dy, real_t dz) {
  LSMHIP_REQUIRE_ARRAYS_OR(-1, phi_out, phi_stage, tabs);
  LSMHIP_REQUIRE_ARRAYS_OR(-1, tabs[0], tabs[1], tabs[2], tabs[3], tabs[4], tabs[5]);
  FusedArgs a = base(phi_out, phi_stage, phi_n, LSMHIP_BOX(gb), LSMHIP_BOX(fb), dt, dx, dy, dz);
  a.rhs = kAdvectDeform;
  for (int d = 0; d < 3; d++) {
    a.ta[d] = tabs[d];
    a.tb[d] = tabs[3 + d];
  }
  a.tfac = c;
  return run("lsmhip_weno5_stage_advect_deform", a, stage);
}

// HJ_WENO5 -> COMPUTE_REINITIALIZATION_EQN_RHS -> RK stage.
LSMHIP_EXPORT int lsmhip_weno5_stage_reinit(real_t* phi_out, const real_t* phi_stage, const real_t* phi_n,
                                            LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), int stage, const real_t* phi0,
                                            int use_phi0_for_sgn, real_t dt, real_t dx, real_t dy, real_t dz) {
  LSMHIP_REQUIRE_ARRAYS_OR(-1, phi_out, phi_stage);
  FusedArgs a = base(phi_out, phi_stage, phi_n, LSMHIP_BOX(gb), LSMHIP_BOX(fb), dt, dx, dy, dz);
  a.rhs = kReinit;
  a.phi0 = phi0;
  a.use_phi0 = use_phi0_for_sgn == 1;
  if (a.use_phi0 && !phi0) {
    set_error(kErrArg, "lsmhip_weno5_stage_reinit: use_phi0_for_sgn set without phi0");
    return -1;
  }
  real_t h = dx > dy ? dx : dy;
  h = h > dz ? h : dz;
  a.dx_sq = h * h;
  return run("lsmhip_weno5_stage_reinit", a, stage);
}

// Writes the deformation velocity of lsmhip_weno5_stage_advect_deform into
// three arrays on gb over box b (for the drop-in chain and checks).
LSMHIP_EXPORT int lsmhip_deform_velocity_fill(real_t* vx, real_t* vy, real_t* vz, LSMHIP_BOXP(gb), LSMHIP_BOXP(b),
                                              const real_t* const* tabs, real_t c) {
  LSMHIP_REQUIRE_ARRAYS_OR(-1, vx, vy, vz, tabs);
  LSMHIP_REQUIRE_ARRAYS_OR(-1, tabs[0], tabs[1], tabs[2], tabs[3], tabs[4], tabs[5]);
  const char* who = "lsmhip_deform_velocity_fill";
  Box gb = LSMHIP_BOX(gb), bb = LSMHIP_BOX(b);
  if (bb.empty()) return 0;
  if (!check_box(gb, bb, who, "vel") || !check_size(gb, who)) return -1;
  hipLaunchKernelGGL(deform_fill_kernel, grid_for(bb), block2d(), 0, stream(), bb, gb, vx, vy, vz, tabs[0], tabs[1],
                     tabs[2], tabs[3], tabs[4], tabs[5], c);
  finish(who);
  return 0;
}

LSMHIP_EXPORT void lsmhip_set_fused_kchunk(int kc) { g_kchunk = kc < 1 ? 1 : kc; }

// FP32 build: 1 (default) selects the stage kernels that form a WENO5 window
// per cell, 2 the kernels that share each position's WENO5 terms between
// neighbouring cells (same scheme, different rounding; slower on gfx950, kept
// for comparisons).  Returns the previous choice; the FP64 build has one set of
// kernels and ignores it.
LSMHIP_EXPORT int lsmhip_set_fused_f32_kernels(int which) {
  const int prev = g_f32_kernels;
  if (which == 1 || which == 2) g_f32_kernels = which;
  return prev;
}
