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

}  // namespace

LSMHIP_EXPORT void lsm3dlinearextrapolation_(real_t* phi, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), const int* idx) {
  face("lsm3dlinearextrapolation_", phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), *idx, kLinear);
}
LSMHIP_EXPORT void lsm3dsignedlinearextrapolation_(real_t* phi, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), const int* idx) {
  face("lsm3dsignedlinearextrapolation_", phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), *idx, kSigned);
}
LSMHIP_EXPORT void lsm3dcopyextrapolation_(real_t* phi, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), const int* idx) {
  face("lsm3dcopyextrapolation_", phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), *idx, kCopy);
}
// Homogeneous Neumann fills of every scheme are copy extrapolation
// (lsm_boundary_conditions3d.f:486-638).
LSMHIP_EXPORT void lsm3dhomogeneousneumanneno1_(real_t* phi, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), const int* idx) {
  face("lsm3dhomogeneousneumanneno1_", phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), *idx, kCopy);
}
LSMHIP_EXPORT void lsm3dhomogeneousneumanneno2_(real_t* phi, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), const int* idx) {
  face("lsm3dhomogeneousneumanneno2_", phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), *idx, kCopy);
}
LSMHIP_EXPORT void lsm3dhomogeneousneumanneno3_(real_t* phi, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), const int* idx) {
  face("lsm3dhomogeneousneumanneno3_", phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), *idx, kCopy);
}
LSMHIP_EXPORT void lsm3dhomogeneousneumannweno5_(real_t* phi, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), const int* idx) {
  face("lsm3dhomogeneousneumannweno5_", phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), *idx, kCopy);
}
// Periodic fill of one face (no reference counterpart: the serial package has
// no periodic boundary condition).  Applied to faces 0..5 in order it fills
// edges and corners periodically too.
LSMHIP_EXPORT void lsmhip_periodic_face_(real_t* phi, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), const int* idx) {
  face("lsmhip_periodic_face_", phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), *idx, kPeriodic);
}

// Dispatch over boundary-location indices 0..9 in the order of the reference's
// C helpers (src/boundary_conditions/lsm_boundary_conditions.c:19-886):
// 0..5 one face, 6 = x faces, 7 = y faces, 8 = z faces, 9 = all six faces in
// the order x_lo, x_hi, y_lo, y_hi, z_lo, z_hi.  kind: 0 linear, 1 signed
// linear, 2 copy (= homogeneous Neumann), 3 periodic.
LSMHIP_EXPORT int lsmhip_fill_ghosts(real_t* phi, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb), int bdry_location_idx, int kind) {
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
  for (int f = first; f <= last; f++) face(who, phi, LSMHIP_BOX(gb), LSMHIP_BOX(fb), f, kind);
  return 0;
}
