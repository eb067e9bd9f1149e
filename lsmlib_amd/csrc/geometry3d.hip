// geometry3d.hip -- the kernels either side of the curvature / reinitialization
// path (whole grid): unit and signed unit normals, mean curvature from a
// precomputed Hessian, Gaussian curvature (second and fourth order),
// fourth-order mean curvature, the orthogonalization RHS and the field-extension
// RHS.
// References: src/geometry/lsm_geometry3d_fort.f.in, src/geometry/lsm_curvature3d.f.in,
// src/reinitialization/lsm_reinitialization3d.f.in, src/field_extension/lsm_field_extension3d.f.
//
// All of them are bound by HBM bandwidth.  A block covers a 64 x 4 tile of
// (i, j) columns and each thread KC consecutive cells along k, tiles dealt to
// the 8 XCDs in contiguous ranges (so in-plane neighbours of a tile are read
// through the same L2).  Pointwise kernels load the inputs of all KC cells
// before computing any (memory-level parallelism); stencil kernels march along
// k carrying the k-neighbours in registers, so each input plane is read from
// HBM once.  In FP64 the operation order is the reference's, without
// contraction, and every result is bitwise the reference's except the mean
// curvature's |grad phi|^3 = (|grad phi|^2)^1.5 (see stencil::pow_three_halves).
#include "lsmhip_core.hpp"
#include "march.hpp"
#include "stencil.hpp"

using namespace lsmhip;

namespace {

bool inside(const char* who, const Box& fb, std::initializer_list<std::pair<Box, const char*>> arrays) {
  for (auto& a : arrays)
    if (!check_box(a.first, fb, who, a.second) || !check_size(a.first, who)) return false;
  return true;
}

// An array read at offsets -w..w along each axis separately (a cross-shaped
// stencil): each widened box must lie inside the array's ghost box.
bool cross_inside(const char* who, const Box& fb, const Box& g, int w, const char* what) {
  return check_box(g, fb.widen(w, 0, 0), who, what) && check_box(g, fb.widen(0, w, 0), who, what) &&
         check_box(g, fb.widen(0, 0, w), who, what) && check_size(g, who);
}

real_t max3(real_t a, real_t b, real_t c) {
  const real_t m = a > b ? a : b;
  return m > c ? m : c;
}

// ---------------------------------------------------------------------------
// Tiling
// ---------------------------------------------------------------------------
struct Tiles {
  Box fb;
  int nti, ntj, ntk, kc, per;
};

Tiles tiles(const Box& fb, int kc) {
  Tiles t{fb, (fb.nx() + kBX - 1) / kBX, (fb.ny() + kBY - 1) / kBY, (fb.nz() + kc - 1) / kc, kc, 0};
  t.per = (int)(((long long)t.nti * t.ntj * t.ntk + 7) / 8);
  return t;
}
dim3 tile_grid(const Tiles& t) { return dim3((unsigned)(8 * t.per)); }

// The column of one thread: cells (i, j, k0 .. k0 + n - 1).  Block b runs on
// XCD b % 8; XCD x takes the tile range [x per, (x + 1) per).
struct Column {
  int i, j, k0, n;
};

__device__ __forceinline__ bool column(const Tiles& T, Column& c) {
  const int b = blockIdx.x;
  const int t = (b % 8) * T.per + b / 8;
  if (t >= T.nti * T.ntj * T.ntk) return false;
  const int ti = t % T.nti, tj = (t / T.nti) % T.ntj, tk = t / (T.nti * T.ntj);
  c.i = T.fb.ilo + ti * kBX + (int)threadIdx.x;
  c.j = T.fb.jlo + tj * kBY + (int)threadIdx.y;
  c.k0 = T.fb.klo + tk * T.kc;
  c.n = min(T.kc, T.fb.khi - c.k0 + 1);
  return c.i <= T.fb.ihi && c.j <= T.fb.jhi;
}

// Pointwise kernels: Op::load(i, j, k) reads one cell's inputs, Op::store
// computes and writes its outputs.  A full column issues all KC loads first.
template <int KC, class Op>
__device__ __forceinline__ void pointwise(const Column& c, const Op& op) {
  if (c.n == KC) {
    typename Op::In v[KC];
#pragma unroll
    for (int t = 0; t < KC; ++t) v[t] = op.load(c.i, c.j, c.k0 + t);
#pragma unroll
    for (int t = 0; t < KC; ++t) op.store(c.i, c.j, c.k0 + t, v[t]);
  } else {
    for (int t = 0; t < c.n; ++t) op.store(c.i, c.j, c.k0 + t, op.load(c.i, c.j, c.k0 + t));
  }
}

template <int KC, class Op>
__global__ void __launch_bounds__(256) pointwise_kernel(Tiles T, Op op) {
  Column c;
  if (!column(T, c)) return;
  pointwise<KC>(c, op);
}

struct Vec3Out {
  View<real_t> x, y, z;
};
struct Vec3In {
  View<const real_t> x, y, z;
};

// ---------------------------------------------------------------------------
// Normals, lsm_geometry3d_fort.f.in:27-272
// ---------------------------------------------------------------------------
struct UnitNormal {
  Vec3Out n;
  Vec3In g;
  struct In {
    real_t x, y, z;
  };
  __device__ __forceinline__ In load(int i, int j, int k) const { return In{g.x(i, j, k), g.y(i, j, k), g.z(i, j, k)}; }
  __device__ __forceinline__ void store(int i, int j, int k, const In& v) const {
    const real_t norm = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    real_t ox = real_t(1), oy = real_t(0), oz = real_t(0);
    if (norm >= kZeroTol) {
      const real_t inv = real_t(1) / norm;
      ox = v.x * inv;
      oy = v.y * inv;
      oz = v.z * inv;
    }
    n.x(i, j, k) = ox;
    n.y(i, j, k) = oy;
    n.z(i, j, k) = oz;
  }
};

struct SignedUnitNormal {
  Vec3Out n;
  Vec3In g;
  View<const real_t> phi;
  real_t dx_sq;
  struct In {
    real_t c, x, y, z;
  };
  __device__ __forceinline__ In load(int i, int j, int k) const {
    return In{phi(i, j, k), g.x(i, j, k), g.y(i, j, k), g.z(i, j, k)};
  }
  __device__ __forceinline__ void store(int i, int j, int k, const In& v) const {
    real_t ox = real_t(0), oy = real_t(0), oz = real_t(0);
    if (fabs(v.c) > kZeroTol) {
      const real_t ngs = v.x * v.x + v.y * v.y + v.z * v.z;
      if (ngs >= kZeroTol) {
        const real_t sgn = v.c / sqrt(v.c * v.c + ngs * dx_sq);
        const real_t inv = real_t(1) / sqrt(ngs);
        ox = sgn * v.x * inv;
        oy = sgn * v.y * inv;
        oz = sgn * v.z * inv;
      } else {
        ox = real_t(1);
      }
    }
    n.x(i, j, k) = ox;
    n.y(i, j, k) = oy;
    n.z(i, j, k) = oz;
  }
};

// ---------------------------------------------------------------------------
// Mean curvature from a given Hessian, lsm_geometry3d_fort.f.in:274-394
// (numerator terms in that routine's order)
// ---------------------------------------------------------------------------
struct MeanCurvatureHessian {
  View<real_t> kappa;
  Vec3In g;
  View<const real_t> hxx, hxy, hxz, hyy, hyz, hzz;
  struct In {
    real_t x, y, z, xx, xy, xz, yy, yz, zz;
  };
  __device__ __forceinline__ In load(int i, int j, int k) const {
    return In{g.x(i, j, k),   g.y(i, j, k),   g.z(i, j, k),   hxx(i, j, k), hxy(i, j, k),
              hxz(i, j, k),   hyy(i, j, k),   hyz(i, j, k),   hzz(i, j, k)};
  }
  __device__ __forceinline__ void store(int i, int j, int k, const In& v) const {
    const real_t x = v.x, y = v.y, z = v.z;
    const real_t ngs = x * x + y * y + z * z;
    real_t out = real_t(0);
    if (!(ngs < kZeroTol)) {
      const real_t num = v.xx * y * y - real_t(2) * v.xy * x * y + v.yy * x * x + v.xx * z * z -
                         real_t(2) * v.xz * x * z + v.zz * x * x + v.yy * z * z - real_t(2) * v.yz * y * z +
                         v.zz * y * y;
      out = num / stencil::pow_three_halves(ngs);
    }
    kappa(i, j, k) = out;
  }
};

// ---------------------------------------------------------------------------
// Orthogonalization RHS, lsm_reinitialization3d.f.in:296-476
// ---------------------------------------------------------------------------
struct Orthogonalization {
  View<real_t> rhs;
  Vec3In phi_p, phi_m;
  View<const real_t> psi;
  Vec3In psi_p, psi_m;
  real_t dx_sq;
  struct In {
    real_t s, qxp, qyp, qzp, qxm, qym, qzm, pxp, pyp, pzp, pxm, pym, pzm;
  };
  __device__ __forceinline__ In load(int i, int j, int k) const {
    return In{psi(i, j, k),     psi_p.x(i, j, k), psi_p.y(i, j, k), psi_p.z(i, j, k), psi_m.x(i, j, k),
              psi_m.y(i, j, k), psi_m.z(i, j, k), phi_p.x(i, j, k), phi_p.y(i, j, k), phi_p.z(i, j, k),
              phi_m.x(i, j, k), phi_m.y(i, j, k), phi_m.z(i, j, k)};
  }
  __device__ __forceinline__ void store(int i, int j, int k, const In& v) const {
    const real_t g1 = real_t(0.5) * (v.qxp + v.qxm);
    const real_t g2 = real_t(0.5) * (v.qyp + v.qym);
    const real_t g3 = real_t(0.5) * (v.qzp + v.qzm);
    const real_t ng = sqrt(g1 * g1 + g2 * g2 + g3 * g3);
    real_t out = real_t(0);
    if (fabs(v.s) > kZeroTol && ng > kZeroTol) {
      const real_t sgn = v.s / sqrt(v.s * v.s + dx_sq * (ng * ng));
      const real_t a1 = sgn * g1, a2 = sgn * g2, a3 = sgn * g3;
      const real_t z = real_t(0);
      out = real_t(-1) / ng *
            (maxr(a1, z) * v.pxm + minr(a1, z) * v.pxp + maxr(a2, z) * v.pym + minr(a2, z) * v.pyp +
             maxr(a3, z) * v.pzm + minr(a3, z) * v.pzp);
    }
    rhs(i, j, k) = out;
  }
};

// ---------------------------------------------------------------------------
// Stencil kernels: march along k with the k-neighbours in registers.
// Pointers are restrict-qualified (the reference's Fortran routines may
// assume that no output aliases an input), so loads of later cells may be
// issued ahead of earlier stores.
// ---------------------------------------------------------------------------
struct Lay {
  int ilo, jlo, klo, sy, sz;
  __device__ __forceinline__ int at(int i, int j, int k) const {
    return (i - ilo) + sy * (j - jlo) + sz * (k - klo);
  }
};
Lay lay(const Box& b) { return Lay{b.ilo, b.jlo, b.klo, b.nx(), b.nx() * b.ny()}; }

// Runs step(t) for t = 0 .. n-1, fully unrolled when the column is full.
template <int KC, class F>
__device__ __forceinline__ void march(int n, F&& step) {
  if (n == KC) {
#pragma unroll
    for (int t = 0; t < KC; ++t) step(t);
  } else {
    for (int t = 0; t < n; ++t) step(t);
  }
}

using stencil::gauss_num;

struct Factors2 {
  real_t xx, yy, zz, xy, xz, yz;
};

// lsm_curvature3d.f.in:136-248: second differences of phi (2-point cross
// stencil for the mixed terms), numerator over |grad phi|^4, on the plane
// march over the 19-point cross; the gradient and its magnitude loaded per
// cell.
struct GaussCurvMarch {
  View<real_t> kappa;
  View<const real_t> px, py, pz, gmag;
  Factors2 f;
  struct In {
    real_t x, y, z, gm;
  };
  __device__ In load(int i, int j, int k) const { return In{px(i, j, k), py(i, j, k), pz(i, j, k), gmag(i, j, k)}; }
  __device__ void store(int i, int j, int k, const lsmhip::march::Cross19& c, const In& v) const {
    const real_t xx = (c.xp - real_t(2) * c.c + c.xm) * f.xx;
    const real_t yy = (c.yp - real_t(2) * c.c + c.ym) * f.yy;
    const real_t zz = (c.zp - real_t(2) * c.c + c.zm) * f.zz;
    const real_t xy = (c.xy[0] - c.xy[1] - c.xy[2] + c.xy[3]) * f.xy;
    const real_t xz = (c.xz[0] - c.xz[1] - c.xz[2] + c.xz[3]) * f.xz;
    const real_t yz = (c.yz[0] - c.yz[1] - c.yz[2] + c.yz[3]) * f.yz;
    const real_t den = v.gm * v.gm * v.gm * v.gm;
    real_t out = real_t(0);
    if (!(den < kZeroTol)) out = gauss_num(v.x, v.y, v.z, xx, yy, zz, xy, xz, yz) / den;
    kappa(i, j, k) = out;
  }
};

struct Factors4 {
  real_t xx, yy, zz;  // 1/(12 h^2)
  real_t x, y, z;     // 1/(12 h)
};

// lsm_curvature3d.f.in:253-507.  Fourth-order second differences of phi
// (:330-341), mixed derivatives as fourth-order differences of the given first
// derivatives (:342-351): phi_xy = d/dy phi_x, phi_yz = d/dz phi_y,
// phi_zx = d/dx phi_z.  GAUSS = 0: mean curvature over |grad phi|^3 (numerator
// order of :358-366); GAUSS = 1: Gaussian curvature over |grad phi|^4.
// On the plane march (64 x 8 tiles, chunks of 16 planes): phi with its x and
// y neighbours, phi_x with its y neighbours and phi_z with its x neighbours
// staged per plane in LDS (two-cell halos), phi and phi_y along z carried in
// registers, every next-plane value in flight while the plane is computed.
// The quotient is formed in every cell and then selected.  g (phi_x, phi_y,
// phi_z, |grad phi|) and phi must cover fb widened by two cells.
template <int GAUSS>
__global__ void __launch_bounds__(lsmhip::march::kMX* lsmhip::march::kMY)
    curvature_order4_march_kernel(Box fb, View<const real_t> F, View<const real_t> gx, View<const real_t> gy,
                                  View<const real_t> gz, View<const real_t> gmag, View<real_t> kappa,
                                  lsmhip::march::Tiles tl, Factors4 f) {
  constexpr int MX = lsmhip::march::kMX, MY = lsmhip::march::kMY, R = 2;
  constexpr int PX = MX + 2 * R, PY = MY + 2 * R;
  constexpr int NP = PX * PY - MX * MY, NX = 2 * R * MX, NZ = 2 * R * MY;
  static_assert(NP <= MX * MY && NX <= MX * MY, "one halo element per thread at most");
  __shared__ real_t P[PX * PY], X[MX * PY], Z[PX * MY];
  const int tid = threadIdx.y * MX + threadIdx.x;
  const int nt = tl.nti * tl.ntj * tl.ntk;
  const int per = (nt + 7) / 8;
  const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (t >= nt) return;
  const int ti = t % tl.nti, tj = (t / tl.nti) % tl.ntj, tk = t / (tl.nti * tl.ntj);
  const int i0 = fb.ilo + ti * MX, j0 = fb.jlo + tj * MY;
  const int kb = fb.klo + tk * tl.kc;
  const int ke = min(kb + tl.kc - 1, fb.khi);
  const int i = i0 + (int)threadIdx.x, j = j0 + (int)threadIdx.y;
  const bool active = i <= fb.ihi && j <= fb.jhi;
  // within two cells of fb: inside phi's and g's boxes; columns further out are never read
  auto near = [&](int a, int b) { return a <= fb.ihi + R && b <= fb.jhi + R; };
  const bool own_in = near(i, j);
  const int psz = F.sz, gsz = gx.sz;
  const int op = own_in ? F.offset(i, j, kb) : 0, og = own_in ? gx.offset(i, j, kb) : 0;
  int hp = -1, hx = -1, hz = -1, sp = 0, sx = 0, szs = 0;
  if (tid < NP) {
    int px, py;
    lsmhip::march::halo_pos<R>(tid, px, py);
    sp = py * PX + px;
    if (near(i0 - R + px, j0 - R + py)) hp = F.offset(i0 - R + px, j0 - R + py, kb);
  }
  if (tid < NX) {  // rows 0, 1 and MY + 2, MY + 3 of the phi_x plane
    const int r = tid / MX, row = r < R ? r : r + MY, px = tid % MX;
    sx = row * MX + px;
    if (near(i0 + px, j0 - R + row)) hx = gx.offset(i0 + px, j0 - R + row, kb);
  }
  if (tid < NZ) {  // columns 0, 1 and MX + 2, MX + 3 of the phi_z plane
    const int row = tid / (2 * R), c = tid % (2 * R), px = c < R ? c : MX + c;
    szs = row * PX + px;
    if (near(i0 - R + px, j0 + row)) hz = gz.offset(i0 - R + px, j0 + row, kb);
  }
  const int slp = (threadIdx.y + R) * PX + threadIdx.x + R, slx = (threadIdx.y + R) * MX + threadIdx.x,
            slz = threadIdx.y * PX + threadIdx.x + R;
  // p[s] = phi(i, j, k - 2 + s), y[s] = phi_y(i, j, k - 2 + s); [5] loading
  real_t p[6], y[6];
#pragma unroll
  for (int s = 0; s < 5; s++) {
    p[s] = own_in ? F.p[op + (s - R) * psz] : real_t(0);
    y[s] = active ? gy.p[og + (s - R) * gsz] : real_t(0);
  }
  real_t cx = own_in ? gx.p[og] : real_t(0), cz = own_in ? gz.p[og] : real_t(0);
  real_t cg = active ? gmag.p[og] : real_t(0);
  real_t vp = hp >= 0 ? F.p[hp] : real_t(0), vx = hx >= 0 ? gx.p[hx] : real_t(0),
         vz = hz >= 0 ? gz.p[hz] : real_t(0);
  for (int k = kb; k <= ke; k++) {
    const int d = k - kb;
    const bool more = k < ke;
    // plane k + 3 of the columns, plane k + 1 of the rest
    p[5] = own_in && more ? F.p[op + (d + 3) * psz] : real_t(0);
    y[5] = active && more ? gy.p[og + (d + 3) * gsz] : real_t(0);
    const real_t ncx = own_in && more ? gx.p[og + (d + 1) * gsz] : real_t(0);
    const real_t ncz = own_in && more ? gz.p[og + (d + 1) * gsz] : real_t(0);
    const real_t ncg = active && more ? gmag.p[og + (d + 1) * gsz] : real_t(0);
    const real_t nvp = hp >= 0 && more ? F.p[hp + (d + 1) * psz] : real_t(0);
    const real_t nvx = hx >= 0 && more ? gx.p[hx + (d + 1) * gsz] : real_t(0);
    const real_t nvz = hz >= 0 && more ? gz.p[hz + (d + 1) * gsz] : real_t(0);
    __syncthreads();  // the previous plane's LDS reads are done
    P[slp] = p[2];
    X[slx] = cx;
    Z[slz] = cz;
    if (tid < NP) P[sp] = vp;
    if (tid < NX) X[sx] = vx;
    if (tid < NZ) Z[szs] = vz;
    __syncthreads();
    if (active) {
      const real_t fxm2 = P[slp - 2], fxm1 = P[slp - 1], fxp1 = P[slp + 1], fxp2 = P[slp + 2];
      const real_t fym2 = P[slp - 2 * PX], fym1 = P[slp - PX], fyp1 = P[slp + PX], fyp2 = P[slp + 2 * PX];
      const real_t xym2 = X[slx - 2 * MX], xym1 = X[slx - MX], xyp1 = X[slx + MX], xyp2 = X[slx + 2 * MX];
      const real_t zxm2 = Z[slz - 2], zxm1 = Z[slz - 1], zxp1 = Z[slz + 1], zxp2 = Z[slz + 2];
      const real_t x = cx, yc = y[2], z = cz, gm = cg;
      const real_t xx = (-fxp2 + real_t(16) * fxp1 - real_t(30) * p[2] - fxm2 + real_t(16) * fxm1) * f.xx;
      const real_t yy = (-fyp2 + real_t(16) * fyp1 - real_t(30) * p[2] - fym2 + real_t(16) * fym1) * f.yy;
      const real_t zz = (-p[4] + real_t(16) * p[3] - real_t(30) * p[2] - p[0] + real_t(16) * p[1]) * f.zz;
      const real_t xy = (-xyp2 + real_t(8) * xyp1 + xym2 - real_t(8) * xym1) * f.y;
      const real_t yz = (-y[4] + real_t(8) * y[3] + y[0] - real_t(8) * y[1]) * f.z;
      const real_t zx = (-zxp2 + real_t(8) * zxp1 + zxm2 - real_t(8) * zxm1) * f.x;
      const real_t den = GAUSS ? gm * gm * gm * gm : gm * gm * gm;
      const real_t num = GAUSS ? gauss_num(x, yc, z, xx, yy, zz, xy, zx, yz)
                               : stencil::curv_num(xx, xy, zx, yy, yz, zz, x, yc, z);
      const real_t quot = num / den;
      kappa(i, j, k) = den < kZeroTol ? real_t(0) : quot;
    }
#pragma unroll
    for (int s = 0; s < 5; s++) {
      p[s] = p[s + 1];
      y[s] = y[s + 1];
    }
    cx = ncx;
    cz = ncz;
    cg = ncg;
    vp = nvp;
    vx = nvx;
    vz = nvz;
  }
}

// ---------------------------------------------------------------------------
// Field extension RHS, lsm_field_extension3d.f:31-156.  Window: phi at k-1,
// k, k+1.
// ---------------------------------------------------------------------------
template <int KC>
__global__ void __launch_bounds__(256)
    field_extension_kernel(Tiles T, real_t* __restrict__ rhs, Lay lr, const real_t* __restrict__ P, Lay lp,
                           const real_t* __restrict__ sx, const real_t* __restrict__ sy_,
                           const real_t* __restrict__ sz_, Lay ls, const real_t* __restrict__ nx,
                           const real_t* __restrict__ ny, const real_t* __restrict__ nz, Lay ln, real_t cutoff) {
  Column c;
  if (!column(T, c)) return;
  const int sy = lp.sy, sz = lp.sz;
  const int op = lp.at(c.i, c.j, c.k0), os = ls.at(c.i, c.j, c.k0), on = ln.at(c.i, c.j, c.k0),
            orr = lr.at(c.i, c.j, c.k0);
  real_t zm = P[op - sz], c0 = P[op], zp = P[op + sz];
  march<KC>(c.n, [&](int t) {
    const int o = op + t * sz;
    const real_t xm = P[o - 1], xp = P[o + 1], ym = P[o - sy], yp = P[o + sy];
    const real_t ax = sx[os + t * ls.sz], ay = sy_[os + t * ls.sz], az = sz_[os + t * ls.sz];
    const real_t bx = nx[on + t * ln.sz], by = ny[on + t * ln.sz], bz = nz[on + t * ln.sz];
    const real_t next = t + 1 < c.n ? P[o + 2 * sz] : zp;
    const bool away = fabs(c0) > cutoff || (c0 * xm > real_t(0) && c0 * xp > real_t(0) && c0 * ym > real_t(0) &&
                                            c0 * yp > real_t(0) && c0 * zm > real_t(0) && c0 * zp > real_t(0));
    rhs[orr + t * lr.sz] = away ? -(bx * ax + by * ay + bz * az) : real_t(0);
    zm = c0;
    c0 = zp;
    zp = next;
  });
}

// ---------------------------------------------------------------------------
// First-order delta function, src/toolbox/lsm_calculus_toolbox3d.f.in:15-141:
// over the six neighbours n where phi changes sign (phi_n phi <= 0), the sum of
// |phi_n N / dx / (phi_n - phi)|, N the normal component of the axis -- except
// that the z terms use norm_phi_x and 1/dx, as the reference's do.
// ---------------------------------------------------------------------------
struct DeltaOrder1 {
  View<real_t> delta;
  View<const real_t> phi, nx, ny;
  real_t ix, iy;
  struct In {
    real_t c, xp, xm, yp, ym, zp, zm, x, y;
  };
  __device__ __forceinline__ In load(int i, int j, int k) const {
    return In{phi(i, j, k),     phi(i + 1, j, k), phi(i - 1, j, k), phi(i, j + 1, k), phi(i, j - 1, k),
              phi(i, j, k + 1), phi(i, j, k - 1), nx(i, j, k),      ny(i, j, k)};
  }
  __device__ __forceinline__ static real_t side(real_t pn, real_t c, real_t n, real_t inv) {
    if (!(pn * c <= real_t(0))) return real_t(0);
    const real_t t = pn * n * inv;
    return t / (pn - c);
  }
  __device__ __forceinline__ void store(int i, int j, int k, const In& v) const {
    delta(i, j, k) = fabs(side(v.xp, v.c, v.x, ix)) + fabs(side(v.xm, v.c, v.x, ix)) +
                     fabs(side(v.yp, v.c, v.y, iy)) + fabs(side(v.ym, v.c, v.y, iy)) +
                     fabs(side(v.zp, v.c, v.x, ix)) + fabs(side(v.zm, v.c, v.x, ix));
  }
};

Vec3In vin(const real_t* x, const real_t* y, const real_t* z, const Box& b) {
  return Vec3In{view(x, b), view(y, b), view(z, b)};
}
Vec3Out vout(real_t* x, real_t* y, real_t* z, const Box& b) { return Vec3Out{view(x, b), view(y, b), view(z, b)}; }

// cells per thread along k
constexpr int kPointKC = 4;   // normals, mean curvature from a Hessian
constexpr int kOrthoKC = 2;   // 13 inputs per cell (9 for the delta function)
constexpr int kMarchKC = 16;  // stencil kernels

template <int KC, class Op>
void launch_pointwise(const Box& fb, const Op& op) {
  const Tiles t = tiles(fb, KC);
  hipLaunchKernelGGL((pointwise_kernel<KC, Op>), tile_grid(t), block2d(), 0, stream(), t, op);
}

}  // namespace

// ---------------------------------------------------------------------------
// Entry points
// ---------------------------------------------------------------------------
LSMHIP_EXPORT void lsm3ddeltafunctionorder1_(const real_t* phi, real_t* delta, LSMHIP_BOXP(pg), const real_t* nx,
                                             const real_t* ny, const real_t* nz, LSMHIP_BOXP(gg), LSMHIP_BOXP(fb),
                                             const real_t* dx, const real_t* dy, const real_t* dz) {
  // norm_phi_z is never read (the reference's z terms use norm_phi_x), so it may be NULL
  LSMHIP_REQUIRE_ARRAYS(phi, delta, nx, ny);
  (void)nz;
  const char* who = "lsm3ddeltafunctionorder1_";
  Box fb = LSMHIP_BOX(fb), pg = LSMHIP_BOX(pg), gg = LSMHIP_BOX(gg);
  if (fb.empty() || !inside(who, fb, {{gg, "norm_phi"}}) || !cross_inside(who, fb, pg, 1, "phi")) return;
  launch_pointwise<kOrthoKC>(fb, DeltaOrder1{view(delta, pg), view(phi, pg), view(nx, gg), view(ny, gg),
                                             real_t(1.0) / *dx, real_t(1.0) / *dy});
  finish(who);
}

LSMHIP_EXPORT void lsm3dcomputeunitnormal_(real_t* nx, real_t* ny, real_t* nz, LSMHIP_BOXP(ng), const real_t* px,
                                           const real_t* py, const real_t* pz, LSMHIP_BOXP(gg), LSMHIP_BOXP(fb)) {
  LSMHIP_REQUIRE_ARRAYS(nx, ny, nz, px, py, pz);
  const char* who = "lsm3dcomputeunitnormal_";
  Box fb = LSMHIP_BOX(fb), ng = LSMHIP_BOX(ng), gg = LSMHIP_BOX(gg);
  if (fb.empty() || !inside(who, fb, {{ng, "normal"}, {gg, "grad_phi"}})) return;
  launch_pointwise<kPointKC>(fb, UnitNormal{vout(nx, ny, nz, ng), vin(px, py, pz, gg)});
  finish(who);
}

LSMHIP_EXPORT void lsm3dcomputesignedunitnormal_(real_t* nx, real_t* ny, real_t* nz, LSMHIP_BOXP(ng),
                                                 const real_t* px, const real_t* py, const real_t* pz,
                                                 LSMHIP_BOXP(gg), const real_t* phi, LSMHIP_BOXP(pg),
                                                 LSMHIP_BOXP(fb), const real_t* dx, const real_t* dy,
                                                 const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(nx, ny, nz, px, py, pz, phi);
  const char* who = "lsm3dcomputesignedunitnormal_";
  Box fb = LSMHIP_BOX(fb), ng = LSMHIP_BOX(ng), gg = LSMHIP_BOX(gg), pg = LSMHIP_BOX(pg);
  if (fb.empty() || !inside(who, fb, {{ng, "normal"}, {gg, "grad_phi"}, {pg, "phi"}})) return;
  const real_t h = max3(*dx, *dy, *dz);
  launch_pointwise<kPointKC>(fb, SignedUnitNormal{vout(nx, ny, nz, ng), vin(px, py, pz, gg), view(phi, pg), h * h});
  finish(who);
}

LSMHIP_EXPORT void lsm3dcomputemeancurvature_(real_t* kappa, LSMHIP_BOXP(kg), const real_t* px, const real_t* py,
                                              const real_t* pz, LSMHIP_BOXP(gg), const real_t* pxx,
                                              const real_t* pxy, const real_t* pxz, const real_t* pyy,
                                              const real_t* pyz, const real_t* pzz, LSMHIP_BOXP(hg),
                                              LSMHIP_BOXP(fb)) {
  LSMHIP_REQUIRE_ARRAYS(kappa, px, py, pz, pxx, pxy, pxz, pyy, pyz, pzz);
  const char* who = "lsm3dcomputemeancurvature_";
  Box fb = LSMHIP_BOX(fb), kg = LSMHIP_BOX(kg), gg = LSMHIP_BOX(gg), hg = LSMHIP_BOX(hg);
  if (fb.empty() || !inside(who, fb, {{kg, "kappa"}, {gg, "grad_phi"}, {hg, "hessian_phi"}})) return;
  launch_pointwise<kPointKC>(fb, MeanCurvatureHessian{view(kappa, kg), vin(px, py, pz, gg), view(pxx, hg),
                                                      view(pxy, hg), view(pxz, hg), view(pyy, hg), view(pyz, hg),
                                                      view(pzz, hg)});
  finish(who);
}

LSMHIP_EXPORT void lsm3dcomputegaussiancurvatureorder2_(real_t* kappa, LSMHIP_BOXP(kg), const real_t* phi,
                                                        LSMHIP_BOXP(pg), const real_t* px, const real_t* py,
                                                        const real_t* pz, const real_t* gmag, LSMHIP_BOXP(gg),
                                                        LSMHIP_BOXP(fb), const real_t* dx, const real_t* dy,
                                                        const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(kappa, phi, px, py, pz, gmag);
  const char* who = "lsm3dcomputegaussiancurvatureorder2_";
  Box fb = LSMHIP_BOX(fb), kg = LSMHIP_BOX(kg), pg = LSMHIP_BOX(pg), gg = LSMHIP_BOX(gg);
  if (fb.empty() || !inside(who, fb, {{kg, "kappa"}, {gg, "grad_phi"}}) ||
      !check_box(pg, fb.widen(1, 1, 1), who, "phi") || !check_size(pg, who))
    return;
  Factors2 f{real_t(1.0) / *dx / *dx, real_t(1.0) / *dy / *dy, real_t(1.0) / *dz / *dz,
             real_t(0.25) / *dx / *dy, real_t(0.25) / *dx / *dz, real_t(0.25) / *dy / *dz};
  const lsmhip::march::Tiles tl = lsmhip::march::tiles_for(fb);
  hipLaunchKernelGGL(lsmhip::march::cross19_march_kernel<GaussCurvMarch>, lsmhip::march::grid_for(tl),
                     lsmhip::march::block(), 0, stream(), fb, view(phi, pg), tl,
                     GaussCurvMarch{view(kappa, kg), view(px, gg), view(py, gg), view(pz, gg), view(gmag, gg), f});
  finish(who);
}

template <int GAUSS>
static void curvature_order4(const char* who, real_t* kappa, const Box& kg, const real_t* phi, const Box& pg,
                             const real_t* px, const real_t* py, const real_t* pz, const real_t* gmag,
                             const Box& gg, const Box& fb, real_t dx, real_t dy, real_t dz) {
  if (fb.empty() || !inside(who, fb, {{kg, "kappa"}}) || !cross_inside(who, fb, gg, 2, "grad_phi") ||
      !cross_inside(who, fb, pg, 2, "phi"))
    return;
  const real_t c = real_t(0.0833333333333333333333);
  Factors4 f{c / dx / dx, c / dy / dy, c / dz / dz, c / dx, c / dy, c / dz};
  const lsmhip::march::Tiles tl = lsmhip::march::tiles_for(fb);
  hipLaunchKernelGGL((curvature_order4_march_kernel<GAUSS>), lsmhip::march::grid_for(tl), lsmhip::march::block(), 0,
                     stream(), fb, view(phi, pg), view(px, gg), view(py, gg), view(pz, gg), view(gmag, gg),
                     view(kappa, kg), tl, f);
  finish(who);
}

LSMHIP_EXPORT void lsm3dcomputemeancurvatureorder4_(real_t* kappa, LSMHIP_BOXP(kg), const real_t* phi,
                                                    LSMHIP_BOXP(pg), const real_t* px, const real_t* py,
                                                    const real_t* pz, const real_t* gmag, LSMHIP_BOXP(gg),
                                                    LSMHIP_BOXP(fb), const real_t* dx, const real_t* dy,
                                                    const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(kappa, phi, px, py, pz, gmag);
  curvature_order4<0>("lsm3dcomputemeancurvatureorder4_", kappa, LSMHIP_BOX(kg), phi, LSMHIP_BOX(pg), px, py, pz,
                      gmag, LSMHIP_BOX(gg), LSMHIP_BOX(fb), *dx, *dy, *dz);
}

LSMHIP_EXPORT void lsm3dcomputegaussiancurvatureorder4_(real_t* kappa, LSMHIP_BOXP(kg), const real_t* phi,
                                                        LSMHIP_BOXP(pg), const real_t* px, const real_t* py,
                                                        const real_t* pz, const real_t* gmag, LSMHIP_BOXP(gg),
                                                        LSMHIP_BOXP(fb), const real_t* dx, const real_t* dy,
                                                        const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(kappa, phi, px, py, pz, gmag);
  curvature_order4<1>("lsm3dcomputegaussiancurvatureorder4_", kappa, LSMHIP_BOX(kg), phi, LSMHIP_BOX(pg), px, py,
                      pz, gmag, LSMHIP_BOX(gg), LSMHIP_BOX(fb), *dx, *dy, *dz);
}

LSMHIP_EXPORT void lsm3dcomputeorthogonalizationeqnrhs_(
    real_t* rhs, LSMHIP_BOXP(rg), const real_t* pxp, const real_t* pyp, const real_t* pzp, LSMHIP_BOXP(gp),
    const real_t* pxm, const real_t* pym, const real_t* pzm, LSMHIP_BOXP(gm), const real_t* psi, LSMHIP_BOXP(sg),
    const real_t* qxp, const real_t* qyp, const real_t* qzp, LSMHIP_BOXP(qp), const real_t* qxm, const real_t* qym,
    const real_t* qzm, LSMHIP_BOXP(qm), LSMHIP_BOXP(fb), const real_t* dx, const real_t* dy, const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(rhs, pxp, pyp, pzp, pxm, pym, pzm, psi, qxp, qyp, qzp, qxm, qym, qzm);
  const char* who = "lsm3dcomputeorthogonalizationeqnrhs_";
  Box fb = LSMHIP_BOX(fb), rg = LSMHIP_BOX(rg), gp = LSMHIP_BOX(gp), gm = LSMHIP_BOX(gm), sg = LSMHIP_BOX(sg),
      qp = LSMHIP_BOX(qp), qm = LSMHIP_BOX(qm);
  if (fb.empty() || !inside(who, fb,
                            {{rg, "ortho_rhs"},
                             {gp, "grad_phi_plus"},
                             {gm, "grad_phi_minus"},
                             {sg, "psi"},
                             {qp, "grad_psi_plus"},
                             {qm, "grad_psi_minus"}}))
    return;
  const real_t h = max3(*dx, *dy, *dz);
  launch_pointwise<kOrthoKC>(fb, Orthogonalization{view(rhs, rg), vin(pxp, pyp, pzp, gp), vin(pxm, pym, pzm, gm),
                                                    view(psi, sg), vin(qxp, qyp, qzp, qp), vin(qxm, qym, qzm, qm),
                                                    h * h});
  finish(who);
}

LSMHIP_EXPORT void lsm3dcomputefieldextensioneqnrhs_(real_t* rhs, LSMHIP_BOXP(rg), const real_t* S, LSMHIP_BOXP(sg),
                                                     const real_t* phi, LSMHIP_BOXP(pg), const real_t* sx,
                                                     const real_t* sy, const real_t* sz, LSMHIP_BOXP(ug),
                                                     const real_t* nx, const real_t* ny, const real_t* nz,
                                                     LSMHIP_BOXP(ng), LSMHIP_BOXP(fb), const real_t* dx,
                                                     const real_t* dy, const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(rhs, S, phi, sx, sy, sz, nx, ny, nz);
  const char* who = "lsm3dcomputefieldextensioneqnrhs_";
  (void)S;  // part of the reference signature; the RHS reads only its upwind gradient
  Box fb = LSMHIP_BOX(fb), rg = LSMHIP_BOX(rg), pg = LSMHIP_BOX(pg), ug = LSMHIP_BOX(ug), ng = LSMHIP_BOX(ng);
  if (fb.empty() || !inside(who, fb, {{rg, "rhs"}, {ug, "grad_S_upwind"}, {ng, "signed_normal"}}) ||
      !cross_inside(who, fb, pg, 1, "phi"))
    return;
  const real_t cutoff = real_t(3.0) * max3(*dx, *dy, *dz);
  const Tiles t = tiles(fb, kMarchKC);
  hipLaunchKernelGGL(field_extension_kernel<kMarchKC>, tile_grid(t), block2d(), 0, stream(), t, rhs, lay(rg), phi,
                     lay(pg), sx, sy, sz, lay(ug), nx, ny, nz, lay(ng), cutoff);
  finish(who);
}
