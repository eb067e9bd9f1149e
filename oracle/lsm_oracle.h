/*
 * lsm_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of LSMLIB's 3D level-set time-stepping kernels, used by
 * tests/, bench.py's cpu_baseline leg and __graft_entry__.smoke() as the
 * checker of the HIP library.  Nothing in the product (lsmlib_amd/) links,
 * loads or calls this code.
 *
 * Every entry point keeps the reference's Fortran ABI (all arguments by
 * pointer, column-major arrays addressed from their own ghost box, fill box
 * and scalars separate) and is exported as `oracle_<fortran symbol>`, e.g.
 * oracle_lsm3dhjweno5 for lsm3dhjweno5_.  Arithmetic is FP64, compiled with
 * -ffp-contract=off, and follows the reference's operation order so that the
 * results are bitwise comparable with the reference kernels.
 */
#ifndef LSM_ORACLE_H
#define LSM_ORACLE_H

#include <math.h>
#include <stdint.h>

/* Template constants of the double-precision build (src/CMakeLists.txt:10-20). */
#define ORC_ZERO_TOL 1.e-11
#define ORC_TINY 1.e-36
#define ORC_BIG 1.e10

/* Six inclusive box bounds passed by pointer, as in every reference signature. */
#define BOXP(n)                                                          \
  const int *ilo_##n, const int *ihi_##n, const int *jlo_##n,            \
      const int *jhi_##n, const int *klo_##n, const int *khi_##n

typedef struct {
  double *p;
  long ilo, jlo, klo, nx, nxy;
} orc_arr;

typedef struct {
  unsigned char *p;
  long ilo, jlo, klo, nx, nxy;
} orc_mask;

static inline orc_arr orc_view(const double *p, const int *ilo, const int *ihi,
                               const int *jlo, const int *jhi, const int *klo) {
  orc_arr a;
  a.p = (double *)p;
  a.ilo = *ilo;
  a.jlo = *jlo;
  a.klo = *klo;
  a.nx = (long)(*ihi - *ilo + 1);
  a.nxy = a.nx * (long)(*jhi - *jlo + 1);
  return a;
}

static inline orc_mask orc_mview(const unsigned char *p, const int *ilo, const int *ihi,
                                 const int *jlo, const int *jhi, const int *klo) {
  orc_mask a;
  a.p = (unsigned char *)p;
  a.ilo = *ilo;
  a.jlo = *jlo;
  a.klo = *klo;
  a.nx = (long)(*ihi - *ilo + 1);
  a.nxy = a.nx * (long)(*jhi - *jlo + 1);
  return a;
}

#define VIEW(p, n) orc_view((p), ilo_##n, ihi_##n, jlo_##n, jhi_##n, klo_##n)
#define MVIEW(p, n) orc_mview((p), ilo_##n, ihi_##n, jlo_##n, jhi_##n, klo_##n)
#define AT(a, i, j, k) \
  ((a).p[((long)(i) - (a).ilo) + (a).nx * ((long)(j) - (a).jlo) + (a).nxy * ((long)(k) - (a).klo)])

/* An inclusive box; 2D boxes have klo = khi = 0. */
typedef struct {
  int ilo, ihi, jlo, jhi, klo, khi;
} obox;

/* Shared between the 3D and 2D restatements (oracle_wholegrid.c): the
 * plus/minus and velocity-upwinded HJ-ENO1/2/3 (order 1..3) and HJ-WENO5
 * (order 5) drivers over the first nd axes, and one face of a ghost fill. */
void orc_hj_plus_minus(int nd, int order, orc_arr *plus, orc_arr *minus, orc_arr phi, orc_arr D1, obox gb1,
                       orc_arr D2, obox gb2, orc_arr D3, obox gb3, obox fb, const double inv[3]);
void orc_hj_upwind(int nd, int order, orc_arr *out, orc_arr phi, orc_arr *vel, orc_arr D1, obox gb1, orc_arr D2,
                   obox gb2, orc_arr D3, obox gb3, obox fb, const double inv[3]);
enum { BC_LINEAR = 0, BC_SIGNED = 1, BC_COPY = 2 };
void orc_bc_face(orc_arr P, obox gb, obox fb, int idx, int kind);

/* Fortran integer*1 marks are compared as signed bytes. */
#define MARK(a, i, j, k) ((signed char)AT(a, i, j, k))

/* Index lists are addressed as the reference's Fortran declarations address
 * them: index_x(nlo:nhi) is passed as a pointer to its element nlo. */
#define LIDX(p, lo, l) ((p)[(l) - (lo)])

/* The band loop of every 3D narrow-band routine: points nlo..nhi of the index
 * lists ix, iy, iz (in scope), those whose mark is <= mark_fb. */
#define FOR_BAND(lo, hi, nbv, mark)                                             \
  for (int l = (lo); l <= (hi); l++)                                            \
    for (int i = LIDX(ix, (lo), l), j = LIDX(iy, (lo), l), k = LIDX(iz, (lo), l), \
             _once = 1;                                                          \
         _once && MARK(nbv, i, j, k) <= (signed char)(mark); _once = 0)

static inline double orc_sq(double x) { return x * x; }
static inline double orc_max(double a, double b) { return a > b ? a : b; }
static inline double orc_min(double a, double b) { return a < b ? a : b; }

#endif
