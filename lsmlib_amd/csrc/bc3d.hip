// bc3d.hip -- ghost-cell fills (src/boundary_conditions/lsm_boundary_conditions3d.f:46-638)
// plus a periodic fill, which the serial reference lacks.
//
// A face kernel maps one thread to one line normal to the face (the two
// transverse indices run over the full ghost box, as in the reference), reads
// the boundary and next-inner values of that line and writes its ghost
// layers.  Lines are independent, so launching faces in the reference's order
// x_lo, x_hi, y_lo, y_hi, z_lo, z_hi reproduces its edge and corner values.
#include "lsmhip_core.hpp"

using namespace lsmhip;

namespace {

enum Kind { kLinear = 0, kSigned = 1, kCopy = 2, kPeriodic = 3 };

__global__ void __launch_bounds__(256) face_kernel(View<real_t> P, Box gb, Box fb, int idx, int kind) {
  const int d = idx / 2, hi = idx % 2;
  const int lo_g[3] = {gb.ilo, gb.jlo, gb.klo}, hi_g[3] = {gb.ihi, gb.jhi, gb.khi};
  const int lo_f[3] = {fb.ilo, fb.jlo, fb.klo}, hi_f[3] = {fb.ihi, fb.jhi, fb.khi};
  const int a = (d + 1) % 3, c = (d + 2) % 3;
  const int na = hi_g[a] - lo_g[a] + 1, nc = hi_g[c] - lo_g[c] + 1;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)na * nc) return;
  const int ta = lo_g[a] + (int)(t % na), tc = lo_g[c] + (int)(t / na);
  const int edge = hi ? hi_f[d] : lo_f[d];
  const int inner = hi ? hi_f[d] - 1 : lo_f[d] + 1;
  const int g0 = hi ? hi_f[d] + 1 : lo_g[d], g1 = hi ? hi_g[d] : lo_f[d] - 1;
  int e[3], n[3];
  e[d] = edge; e[a] = ta; e[c] = tc;
  n[d] = inner; n[a] = ta; n[c] = tc;
  if (kind == kPeriodic) {
    const int period = hi_f[d] - lo_f[d] + 1;
    for (int g = g0; g <= g1; g++) {
      int q[3], s[3];
      q[d] = g; q[a] = ta; q[c] = tc;
      s[d] = hi ? g - period : g + period; s[a] = ta; s[c] = tc;
      P(q[0], q[1], q[2]) = P(s[0], s[1], s[2]);
    }
    return;
  }
  const real_t pb = P(e[0], e[1], e[2]);
  real_t slope = real_t(0);
  if (kind == kLinear) slope = pb - P(n[0], n[1], n[2]);
  if (kind == kSigned) slope = copysign(real_t(1), pb) * fabs(pb - P(n[0], n[1], n[2]));
  for (int g = g0; g <= g1; g++) {
    int q[3];
    q[d] = g; q[a] = ta; q[c] = tc;
    if (kind == kCopy) {
      P(q[0], q[1], q[2]) = pb;
    } else {
      const real_t dist = real_t(hi ? g - edge : edge - g);
      P(q[0], q[1], q[2]) = pb + slope * dist;
    }
  }
}

void face(const char* who, real_t* phi, const Box& gb, const Box& fb, int idx, int kind) {
  if (idx < 0 || idx > 5) return;  // the reference ignores other indices
  if (gb.empty() || fb.empty()) return;
  if (!check_box(gb, fb, who, "phi") || !check_size(gb, who)) return;
  const int d = idx / 2;
  const int lo_f[3] = {fb.ilo, fb.jlo, fb.klo}, hi_f[3] = {fb.ihi, fb.jhi, fb.khi};
  const int lo_g[3] = {gb.ilo, gb.jlo, gb.klo}, hi_g[3] = {gb.ihi, gb.jhi, gb.khi};
  const int depth = idx % 2 ? hi_g[d] - hi_f[d] : lo_f[d] - lo_g[d];
  if (depth <= 0) return;
  if (kind == kLinear || kind == kSigned) {
    const int inner = idx % 2 ? hi_f[d] - 1 : lo_f[d] + 1;  // the slope reads this layer
    if (inner < lo_g[d] || inner > hi_g[d]) {
      set_error(kErrBox, "%s: the layer next to face %d lies outside the ghost box", who, idx);
      return;
    }
  }
  if (kind == kPeriodic && hi_f[d] - lo_f[d] + 1 < depth) {
    set_error(kErrBox, "%s: periodic fill needs at least %d interior cells normal to face %d", who, depth, idx);
    return;
  }
  const int a = (d + 1) % 3, c = (d + 2) % 3;
  const long long lines = (long long)(hi_g[a] - lo_g[a] + 1) * (hi_g[c] - lo_g[c] + 1);
  hipLaunchKernelGGL(face_kernel, dim3((unsigned)((lines + 255) / 256)), dim3(256), 0, stream(), view(phi, gb), gb,
                     fb, idx, kind);
  finish(who);
}

// All six faces in one launch.  Applying faces x_lo, x_hi, y_lo, y_hi, z_lo,
// z_hi in turn leaves in every ghost cell the value written by the face of its
// last ghost axis (z before y before x), computed from cells whose coordinate
// on that axis is the edge or next-inner layer; those cells' own values come
// from the faces of the remaining ghost axes, all applied earlier.  So the final
// value of a ghost cell is that recursion evaluated from interior cells only,
// with the face kernel's operations, and no ghost cell is read.
//
// The launch covers the ghost shell by lines normal to a face: one thread per
// (i, j) column of the ghost box for the k ghost layers, one per (i, k) with
// k interior for the j ghost layers, and eight lanes per (j, k) row with j and
// k interior for the i ghost layers (adjacent lanes write adjacent cells, so a
// row's few ghost cells go out in one store request).  A line reads its edge
// and next-inner values once and writes every ghost layer on that side from
// them.
struct Shell {
  Box gb, fb;
  long long lz, ly, lx;  // lines normal to the z, y and x faces
  int kind;
};

__device__ __forceinline__ real_t extrap(real_t pb, real_t pin, int dist, int kind) {
  if (kind == kCopy) return pb;
  const real_t slope = kind == kLinear ? pb - pin : copysign(real_t(1), pb) * fabs(pb - pin);
  return pb + slope * real_t(dist);
}

// Value of cell (i, j, k) once the faces of axis 0 (v0: j, k interior) and of
// axes 0..1 (v1: k interior) have been applied.
__device__ __forceinline__ real_t shell_v0(const View<real_t>& P, const Shell& S, int i, int j, int k) {
  const Box& f = S.fb;
  if (i >= f.ilo && i <= f.ihi) return P(i, j, k);
  const bool up = i > f.ihi;
  if (S.kind == kPeriodic) return P(up ? i - f.nx() : i + f.nx(), j, k);
  const int e = up ? f.ihi : f.ilo, n = up ? f.ihi - 1 : f.ilo + 1;
  return extrap(P(e, j, k), S.kind == kCopy ? real_t(0) : P(n, j, k), up ? i - e : e - i, S.kind);
}
__device__ __forceinline__ real_t shell_v1(const View<real_t>& P, const Shell& S, int i, int j, int k) {
  const Box& f = S.fb;
  if (j >= f.jlo && j <= f.jhi) return shell_v0(P, S, i, j, k);
  const bool up = j > f.jhi;
  if (S.kind == kPeriodic) return shell_v0(P, S, i, up ? j - f.ny() : j + f.ny(), k);
  const int e = up ? f.jhi : f.jlo, n = up ? f.jhi - 1 : f.jlo + 1;
  const real_t pb = shell_v0(P, S, i, e, k);
  return extrap(pb, S.kind == kCopy ? real_t(0) : shell_v0(P, S, i, n, k), up ? j - e : e - j, S.kind);
}

__global__ void __launch_bounds__(256) shell_kernel(View<real_t> P, Shell S) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const Box &g = S.gb, &f = S.fb;
  if (t < S.lz) {  // column (i, j): the k ghost layers
    const int nxg = g.nx(), c = (int)t;
    const int i = g.ilo + c % nxg, j = g.jlo + c / nxg;
    for (int up = 0; up < 2; up++) {
      const int k0 = up ? f.khi + 1 : g.klo, k1 = up ? g.khi : f.klo - 1;
      if (S.kind == kPeriodic) {
        for (int k = k0; k <= k1; k++) P(i, j, k) = shell_v1(P, S, i, j, up ? k - f.nz() : k + f.nz());
      } else if (k0 <= k1) {
        const int e = up ? f.khi : f.klo, n = up ? f.khi - 1 : f.klo + 1;
        const real_t pb = shell_v1(P, S, i, j, e);
        const real_t pin = S.kind == kCopy ? real_t(0) : shell_v1(P, S, i, j, n);
        for (int k = k0; k <= k1; k++) P(i, j, k) = extrap(pb, pin, up ? k - e : e - k, S.kind);
      }
    }
  } else if (t < S.lz + S.ly) {  // row (i, k), k interior: the j ghost layers
    const int nxg = g.nx(), c = (int)(t - S.lz);
    const int i = g.ilo + c % nxg, k = f.klo + c / nxg;
    for (int up = 0; up < 2; up++) {
      const int j0 = up ? f.jhi + 1 : g.jlo, j1 = up ? g.jhi : f.jlo - 1;
      if (S.kind == kPeriodic) {
        for (int j = j0; j <= j1; j++) P(i, j, k) = shell_v0(P, S, i, up ? j - f.ny() : j + f.ny(), k);
      } else if (j0 <= j1) {
        const int e = up ? f.jhi : f.jlo, n = up ? f.jhi - 1 : f.jlo + 1;
        const real_t pb = shell_v0(P, S, i, e, k);
        const real_t pin = S.kind == kCopy ? real_t(0) : shell_v0(P, S, i, n, k);
        for (int j = j0; j <= j1; j++) P(i, j, k) = extrap(pb, pin, up ? j - e : e - j, S.kind);
      }
    }
  } else {  // row (j, k), j and k interior: the i ghost layers, eight lanes per row
    const long long u = t - S.lz - S.ly;
    if (u >= 8 * S.lx) return;
    const int r = (int)(u >> 3), q = (int)(u & 7), nyi = f.ny();
    const int j = f.jlo + r % nyi, k = f.klo + r / nyi;
    const int below = f.ilo - g.ilo, depth = below + (g.ihi - f.ihi);
    for (int m = q; m < depth; m += 8) {
      const bool up = m >= below;
      const int i = up ? f.ihi + 1 + (m - below) : g.ilo + m;
      if (S.kind == kPeriodic) {
        P(i, j, k) = P(up ? i - f.nx() : i + f.nx(), j, k);
      } else {
        const int e = up ? f.ihi : f.ilo, n = up ? f.ihi - 1 : f.ilo + 1;
        P(i, j, k) = extrap(P(e, j, k), S.kind == kCopy ? real_t(0) : P(n, j, k), up ? i - e : e - i, S.kind);
      }
    }
  }
}

// Checks of face() for all six faces, then one shell launch.
int shell(const char* who, real_t* phi, const Box& gb, const Box& fb, int kind) {
  if (gb.empty() || fb.empty()) return 0;
  if (!check_box(gb, fb, who, "phi") || !check_size(gb, who)) return -1;
  const int lo_f[3] = {fb.ilo, fb.jlo, fb.klo}, hi_f[3] = {fb.ihi, fb.jhi, fb.khi};
  const int lo_g[3] = {gb.ilo, gb.jlo, gb.klo}, hi_g[3] = {gb.ihi, gb.jhi, gb.khi};
  for (int idx = 0; idx < 6; idx++) {
    const int d = idx / 2;
    const int depth = idx % 2 ? hi_g[d] - hi_f[d] : lo_f[d] - lo_g[d];
    if (depth <= 0) continue;
    if ((kind == kLinear || kind == kSigned) && hi_f[d] == lo_f[d]) {
      set_error(kErrBox, "%s: the layer next to face %d lies outside the ghost box", who, idx);
      return -1;
    }
    if (kind == kPeriodic && hi_f[d] - lo_f[d] + 1 < depth) {
      set_error(kErrBox, "%s: periodic fill needs at least %d interior cells normal to face %d", who, depth, idx);
      return -1;
    }
  }
  // lines exist only normal to faces with ghost layers
  const bool gz = fb.klo > gb.klo || fb.khi < gb.khi, gy = fb.jlo > gb.jlo || fb.jhi < gb.jhi;
  const bool gx = fb.ilo > gb.ilo || fb.ihi < gb.ihi;
  Shell S{gb, fb, 0, 0, 0, kind};
  S.lz = gz ? (long long)gb.nx() * gb.ny() : 0;
  S.ly = gy ? (long long)gb.nx() * fb.nz() : 0;
  S.lx = gx ? (long long)fb.ny() * fb.nz() : 0;
  const long long threads = S.lz + S.ly + 8 * S.lx;
  if (threads == 0) return 0;
  hipLaunchKernelGGL(shell_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, stream(), view(phi, gb), S);
  return finish(who) ? 0 : -1;
}

}  // namespace

LSMHIP_EXPORT void lsm3dlinearextrapolation_(real_t* phi, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), const int* idx) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  face("lsm3dlinearextrapolation_", phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), *idx, kLinear);
}
LSMHIP_EXPORT void lsm3dsignedlinearextrapolation_(real_t* phi, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), const int* idx) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  face("lsm3dsignedlinearextrapolation_", phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), *idx, kSigned);
}
LSMHIP_EXPORT void lsm3dcopyextrapolation_(real_t* phi, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), const int* idx) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  face("lsm3dcopyextrapolation_", phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), *idx, kCopy);
}
// Homogeneous Neumann fills of every scheme are copy extrapolation
// (lsm_boundary_conditions3d.f:486-638).
LSMHIP_EXPORT void lsm3dhomogeneousneumanneno1_(real_t* phi, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), const int* idx) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  face("lsm3dhomogeneousneumanneno1_", phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), *idx, kCopy);
}
LSMHIP_EXPORT void lsm3dhomogeneousneumanneno2_(real_t* phi, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), const int* idx) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  face("lsm3dhomogeneousneumanneno2_", phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), *idx, kCopy);
}
LSMHIP_EXPORT void lsm3dhomogeneousneumanneno3_(real_t* phi, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), const int* idx) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  face("lsm3dhomogeneousneumanneno3_", phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), *idx, kCopy);
}
LSMHIP_EXPORT void lsm3dhomogeneousneumannweno5_(real_t* phi, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), const int* idx) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  face("lsm3dhomogeneousneumannweno5_", phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), *idx, kCopy);
}
// Periodic fill of one face (no reference counterpart: the serial package has
// no periodic boundary condition).  Applied to faces 0..5 in order it fills
// edges and corners periodically too.
LSMHIP_EXPORT void lsmhip_periodic_face_(real_t* phi, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), const int* idx) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  face("lsmhip_periodic_face_", phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), *idx, kPeriodic);
}

// Dispatch over boundary-location indices 0..9 in the order of the reference's
// C helpers (src/boundary_conditions/lsm_boundary_conditions.c:19-886):
// 0..5 one face, 6 = x faces, 7 = y faces, 8 = z faces, 9 = all six faces in
// the order x_lo, x_hi, y_lo, y_hi, z_lo, z_hi.  kind: 0 linear, 1 signed
// linear, 2 copy (= homogeneous Neumann), 3 periodic.
LSMHIP_EXPORT int lsmhip_fill_ghosts(real_t* phi, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), int bdry_location_idx, int kind) {
  LSMHIP_REQUIRE_ARRAYS_OR(-1, phi);
  static const char* who = "lsmhip_fill_ghosts";
  if (kind < 0 || kind > 3) {
    set_error(kErrArg, "%s: unknown fill kind %d", who, kind);
    return -1;
  }
  int first, last;
  if (bdry_location_idx >= 0 && bdry_location_idx <= 5) first = last = bdry_location_idx;
  else if (bdry_location_idx >= 6 && bdry_location_idx <= 8) {
    first = 2 * (bdry_location_idx - 6);
    last = first + 1;
  } else if (bdry_location_idx == 9) {
    first = 0;
    last = 5;
  } else {
    set_error(kErrArg, "%s: boundary location index %d not in 0..9", who, bdry_location_idx);
    return -1;
  }
  if (bdry_location_idx == 9) return shell(who, phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), kind);
  for (int f = first; f <= last; f++) face(who, phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), f, kind);
  return 0;
}

// ---------------------------------------------------------------------------
// The reference's ghost-fill dispatch over a Grid, under its own names
// (src/boundary_conditions/lsm_boundary_conditions.c:19-886).  A 3D grid takes
// lsmhip_fill_ghosts on the grid's ghost and fill boxes (the same face groups
// and order); a 2D grid takes the 2D face fills one after the other.  Location
// indices the reference's switch does not list do nothing, as there.
// ---------------------------------------------------------------------------
namespace {
void grid_fill(const char* who, real_t* phi, const Grid* g, int idx, int kind) {
  if (!g) {
    set_error(kErrArg, "%s: grid is NULL (nothing was done)", who);
    return;
  }
  if (!phi) {
    set_error(kErrArg, "%s: array argument phi is NULL (nothing was done)", who);
    return;
  }
  if (g->num_dims == 3) {
    if (idx >= 0 && idx <= 9)
      lsmhip_fill_ghosts(phi, &g->ilo_gb, &g->ihi_gb, &g->jlo_gb, &g->jhi_gb, &g->klo_gb, &g->khi_gb, &g->ilo_fb,
                         &g->ihi_fb, &g->jlo_fb, &g->jhi_fb, &g->klo_fb, &g->khi_fb, idx, kind);
    return;
  }
  if (g->num_dims != 2) return;
  // 2D: faces 0..3, groups 6 (x), 7 (y) and 9 (all four); 4, 5 and 8 are not in the reference's 2D switch
  int first = 0, last = -1;
  if (idx >= 0 && idx <= 3) first = last = idx;
  else if (idx == 6 || idx == 7) {
    first = 2 * (idx - 6);
    last = first + 1;
  } else if (idx == 9) {
    first = 0;
    last = 3;
  }
  void (*const fill[3])(real_t*, const int*, const int*, const int*, const int*, const int*, const int*, const int*,
                        const int*, const int*) = {lsm2dlinearextrapolation_, lsm2dsignedlinearextrapolation_,
                                                   lsm2dcopyextrapolation_};
  for (int f = first; f <= last; f++)
    fill[kind](phi, &g->ilo_gb, &g->ihi_gb, &g->jlo_gb, &g->jhi_gb, &g->ilo_fb, &g->ihi_fb, &g->jlo_fb, &g->jhi_fb,
               &f);
}
}  // namespace

LSMHIP_EXPORT void linearExtrapolationBC(real_t* phi, Grid* grid, int bdry_location_idx) {
  grid_fill("linearExtrapolationBC", phi, grid, bdry_location_idx, kLinear);
}
LSMHIP_EXPORT void signedLinearExtrapolationBC(real_t* phi, Grid* grid, int bdry_location_idx) {
  grid_fill("signedLinearExtrapolationBC", phi, grid, bdry_location_idx, kSigned);
}
LSMHIP_EXPORT void copyExtrapolationBC(real_t* phi, Grid* grid, int bdry_location_idx) {
  grid_fill("copyExtrapolationBC", phi, grid, bdry_location_idx, kCopy);
}
// homogeneousNeumannBC is copy extrapolation (lsm_boundary_conditions.c:880-886)
LSMHIP_EXPORT void homogeneousNeumannBC(real_t* phi, Grid* grid, int bdry_location_idx) {
  grid_fill("homogeneousNeumannBC", phi, grid, bdry_location_idx, kCopy);
}
