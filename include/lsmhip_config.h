/*
 * lsmhip_config.h -- precision and argument conventions of the HIP drop-in
 * for LSMLIB's level-set time-stepping kernels (3D, and their 2D twins).
 *
 * liblsmhip.so is built with LSMLIB_REAL = double and liblsmhip_f32.so with
 * LSMLIB_REAL = float; both export the same symbols, as the reference's double
 * and single-precision builds do (reference include/lsmlib_config.h.in:24,
 * CMakeLists.txt:36-41).  Define LSMLIB_REAL as float before including these
 * headers to call the FP32 library.
 *
 * Conventions kept from the reference (src/toolbox/lsm_spatial_derivatives3d.h:36-50
 * and every other header of the path): all arguments are pointers; arrays are
 * column-major, each addressed from the first element of its own ghost box;
 * boxes are six ints (ilo, ihi, jlo, jhi, klo, khi), four (ilo, ihi, jlo, jhi)
 * for the 2D twins.  Different from the
 * reference: array arguments are device pointers (hipMalloc'd or torch
 * tensors on the GPU); scalars, boxes, n_lo/n_hi level tables and scalar
 * outputs stay host pointers.  Calls complete before returning unless
 * lsmhip_set_stream(stream, 0) selected stream-ordered mode.
 */
#ifndef LSMHIP_CONFIG_H
#define LSMHIP_CONFIG_H

#ifndef LSMLIB_REAL
#define LSMLIB_REAL double
#endif

/* Six inclusive box bounds of an array (its ghost box) or of a fill box. */
#define LSMHIP_BOX_ARGS(n)                                                              \
  const int *ilo_##n, const int *ihi_##n, const int *jlo_##n, const int *jhi_##n,      \
      const int *klo_##n, const int *khi_##n

/* Four inclusive box bounds of a 2D array or fill box (ilo, ihi, jlo, jhi). */
#define LSMHIP_BOX2_ARGS(n) const int *ilo_##n, const int *ihi_##n, const int *jlo_##n, const int *jhi_##n

#ifdef __cplusplus
#define LSMHIP_BEGIN_DECLS extern "C" {
#define LSMHIP_END_DECLS }
#else
#define LSMHIP_BEGIN_DECLS
#define LSMHIP_END_DECLS
#endif

#endif
