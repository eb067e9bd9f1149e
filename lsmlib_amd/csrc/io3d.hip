// io3d.hip -- the reference's binary file formats for level-set data arrays and
// grids, so that fields computed on the GPU can be read by the reference's
// tools (and its MATLAB readers) and the other way round.
//
//   data array  (src/utils/lsm_data_arrays.c:364-437): three native ints, the
//               ghost-box dimensions, then the reals of the array, column-major;
//   grid        (src/utils/lsm_grid.c:614-763): the Grid fields in declaration
//               order, each written with its own size (no padding).
//
// The Grid construction (src/utils/lsm_grid.c:47-232, 789-1031) is restated so
// that a grid built here carries the same ghost width, fill boxes, marks and
// band widths as one built by the reference.  Host code only; data arrays may
// live in device memory (copied through a host buffer) or on the host.
// Compression (the reference's zip_status, which shells out to gzip) is left
// to the caller.
#include <cmath>
#include <cstdio>
#include <vector>

#include "lsmhip_core.hpp"

using namespace lsmhip;

namespace {

// number of ghost cells per accuracy type, src/utils/lsm_grid.c:47
const int kGhost[4] = {2, 3, 5, 4};

struct File {
  FILE* f;
  explicit File(FILE* p) : f(p) {}
  ~File() {
    if (f) fclose(f);
  }
};

template <class T>
bool put(FILE* f, const T* v, size_t n) {
  return fwrite(v, sizeof(T), n, f) == n;
}
template <class T>
bool get(FILE* f, T* v, size_t n) {
  return fread(v, sizeof(T), n, f) == n;
}

// Copy between a caller array (device or host) and a host buffer in chunks.
bool to_host(real_t* dst, const real_t* src, long long n, int on_device, const char* who) {
  if (!on_device) {
    std::copy(src, src + n, dst);
    return true;
  }
  return check_hip(hipMemcpy(dst, src, sizeof(real_t) * n, hipMemcpyDeviceToHost), who);
}
bool from_host(real_t* dst, const real_t* src, long long n, int on_device, const char* who) {
  if (!on_device) {
    std::copy(src, src + n, dst);
    return true;
  }
  return check_hip(hipMemcpy(dst, src, sizeof(real_t) * n, hipMemcpyHostToDevice), who);
}

constexpr long long kChunk = 1 << 24;  // elements per host staging buffer

void index_limits_central(LsmhipGrid* g) {  // src/utils/lsm_grid.c:789-808
  const int* n = g->grid_dims_ghostbox;
  g->ilo_gb = 0; g->ihi_gb = n[0] - 1; g->jlo_gb = 0; g->jhi_gb = n[1] - 1; g->klo_gb = 0; g->khi_gb = n[2] - 1;
  g->ilo_D1_fb = 1; g->ihi_D1_fb = n[0] - 2; g->jlo_D1_fb = 1; g->jhi_D1_fb = n[1] - 2;
  g->klo_D1_fb = 1; g->khi_D1_fb = n[2] - 2;
  g->ilo_D2_fb = 2; g->ihi_D2_fb = n[0] - 3; g->jlo_D2_fb = 2; g->jhi_D2_fb = n[1] - 3;
  g->klo_D2_fb = 2; g->khi_D2_fb = n[2] - 3;
  g->ilo_D3_fb = 3; g->ihi_D3_fb = n[0] - 4; g->jlo_D3_fb = 3; g->jhi_D3_fb = n[1] - 4;
  g->klo_D3_fb = 3; g->khi_D3_fb = n[2] - 4;
}

// The common fill box of the plus and minus derivatives of each scheme,
// src/utils/lsm_grid.c:832-985: minus starts `lo` cells into its difference
// box, plus ends `hi` cells before it.
void fill_box(LsmhipGrid* g, const int* D_lo, const int* D_hi, int lo, int hi) {
  g->ilo_fb = D_lo[0] + lo; g->ihi_fb = D_hi[0] - hi;
  g->jlo_fb = D_lo[1] + lo; g->jhi_fb = D_hi[1] - hi;
  g->klo_fb = D_lo[2] + lo; g->khi_fb = D_hi[2] - hi;
}

void index_limits(int accuracy, LsmhipGrid* g) {  // src/utils/lsm_grid.c:987-1031
  g->mark_gb = 120;
  g->mark_D1 = (unsigned char)(g->mark_gb - 1);
  g->mark_D2 = (unsigned char)(g->mark_gb - 2);
  g->mark_D3 = (unsigned char)(g->mark_gb - 3);
  g->num_nb_levels = kGhost[accuracy];
  g->mark_fb = (unsigned char)(g->mark_gb - g->num_nb_levels);
  if (accuracy > 1) {
    g->beta = 3 * g->dx[0];
    g->gamma = 6 * g->dx[0];
  } else {
    g->beta = 2 * g->dx[0];
    g->gamma = 4 * g->dx[0];
  }
  index_limits_central(g);
  const int D1lo[3] = {g->ilo_D1_fb, g->jlo_D1_fb, g->klo_D1_fb}, D1hi[3] = {g->ihi_D1_fb, g->jhi_D1_fb, g->khi_D1_fb};
  const int D2lo[3] = {g->ilo_D2_fb, g->jlo_D2_fb, g->klo_D2_fb}, D2hi[3] = {g->ihi_D2_fb, g->jhi_D2_fb, g->khi_D2_fb};
  const int D3lo[3] = {g->ilo_D3_fb, g->jlo_D3_fb, g->klo_D3_fb}, D3hi[3] = {g->ihi_D3_fb, g->jhi_D3_fb, g->khi_D3_fb};
  switch (accuracy) {
    case 0: fill_box(g, D1lo, D1hi, 0, 1); break;   // ENO1
    case 2: fill_box(g, D3lo, D3hi, 1, 2); break;   // ENO3
    case 3: fill_box(g, D1lo, D1hi, 2, 3); break;   // WENO5
    default: fill_box(g, D2lo, D2hi, 1, 1); break;  // ENO2
  }
}

void finish_dims(LsmhipGrid* g, int accuracy) {  // src/utils/lsm_grid.c:80-105, 218-232
  const int ng = kGhost[accuracy];
  g->num_gridpts = 1;
  for (int i = 0; i < g->num_dims; i++) {
    g->grid_dims_ghostbox[i] = g->grid_dims[i] + 2 * ng;
    g->x_lo_ghostbox[i] = g->x_lo[i] - ng * g->dx[i];
    g->x_hi_ghostbox[i] = g->x_hi[i] + ng * g->dx[i];
    g->num_gridpts *= g->grid_dims_ghostbox[i];
  }
  if (g->num_dims == 2) {
    g->x_lo[2] = 0; g->x_hi[2] = 0; g->x_lo_ghostbox[2] = 0; g->x_hi_ghostbox[2] = 0;
    g->grid_dims[2] = 1; g->grid_dims_ghostbox[2] = 1; g->dx[2] = 0;
  }
  index_limits(accuracy, g);
}

bool check_grid_args(int num_dims, int accuracy, const char* who) {
  if (num_dims != 2 && num_dims != 3) {
    set_error(kErrArg, "%s: num_dims must be 2 or 3, got %d", who, num_dims);
    return false;
  }
  if (accuracy < 0 || accuracy > 3) {
    set_error(kErrArg, "%s: accuracy must be 0 (ENO1) .. 3 (WENO5), got %d", who, accuracy);
    return false;
  }
  return true;
}

}  // namespace

// writeDataArrayNoGrid / writeDataArray without compression
// (src/utils/lsm_data_arrays.c:364-399).  data: grid_dims_ghostbox[0..2]
// column-major reals, in device memory when on_device != 0.
LSMHIP_EXPORT int lsmhip_write_data_array(const real_t* data, const int* grid_dims_ghostbox, const char* file_name,
                                          int on_device) {
  const char* who = "lsmhip_write_data_array";
  const long long n = (long long)grid_dims_ghostbox[0] * grid_dims_ghostbox[1] * grid_dims_ghostbox[2];
  File f(fopen(file_name, "wb"));
  if (!f.f) {
    set_error(kErrArg, "%s: cannot open %s for writing", who, file_name);
    return -1;
  }
  if (!put(f.f, grid_dims_ghostbox, 3)) {
    set_error(kErrArg, "%s: short write to %s", who, file_name);
    return -1;
  }
  std::vector<real_t> buf((size_t)std::min(n, kChunk));
  for (long long o = 0; o < n; o += kChunk) {
    const long long m = std::min(kChunk, n - o);
    if (!to_host(buf.data(), data + o, m, on_device, who)) return -1;
    if (!put(f.f, buf.data(), (size_t)m)) {
      set_error(kErrArg, "%s: short write to %s", who, file_name);
      return -1;
    }
  }
  return 0;
}

// The ghost-box dimensions stored in a data-array file.
LSMHIP_EXPORT int lsmhip_read_data_array_dims(int* grid_dims_ghostbox, const char* file_name) {
  const char* who = "lsmhip_read_data_array_dims";
  File f(fopen(file_name, "rb"));
  if (!f.f || !get(f.f, grid_dims_ghostbox, 3)) {
    set_error(kErrArg, "%s: cannot read the header of %s", who, file_name);
    return -1;
  }
  return 0;
}

// readDataArray (src/utils/lsm_data_arrays.c:402-437) into a caller array of
// `capacity` reals (device memory when on_device != 0); the file's dimensions
// are returned in grid_dims_ghostbox.  A file larger than capacity or shorter
// than its header says is refused.
LSMHIP_EXPORT int lsmhip_read_data_array(real_t* data, long long capacity, int* grid_dims_ghostbox,
                                         const char* file_name, int on_device) {
  const char* who = "lsmhip_read_data_array";
  File f(fopen(file_name, "rb"));
  if (!f.f || !get(f.f, grid_dims_ghostbox, 3)) {
    set_error(kErrArg, "%s: cannot read the header of %s", who, file_name);
    return -1;
  }
  const long long n = (long long)grid_dims_ghostbox[0] * grid_dims_ghostbox[1] * grid_dims_ghostbox[2];
  if (n < 0 || n > capacity) {
    set_error(kErrArg, "%s: %s holds %lld reals, the array takes %lld", who, file_name, n, capacity);
    return -1;
  }
  std::vector<real_t> buf((size_t)std::min(n, kChunk));
  for (long long o = 0; o < n; o += kChunk) {
    const long long m = std::min(kChunk, n - o);
    if (!get(f.f, buf.data(), (size_t)m)) {
      set_error(kErrArg, "%s: %s ends before its %lld reals", who, file_name, n);
      return -1;
    }
    if (!from_host(data + o, buf.data(), m, on_device, who)) return -1;
  }
  return 0;
}

// createGridSetGridDims (src/utils/lsm_grid.c:189-232): accuracy 0 ENO1, 1 ENO2,
// 2 ENO3, 3 WENO5.
LSMHIP_EXPORT int lsmhip_grid_set_grid_dims(LsmhipGrid* g, int num_dims, const int* grid_dims, const real_t* x_lo,
                                            const real_t* x_hi, int accuracy) {
  if (!check_grid_args(num_dims, accuracy, "lsmhip_grid_set_grid_dims")) return -1;
  *g = LsmhipGrid{};
  g->num_dims = num_dims;
  for (int i = 0; i < num_dims; i++) {
    g->grid_dims[i] = grid_dims[i];
    g->x_hi[i] = x_hi[i];
    g->x_lo[i] = x_lo[i];
    g->dx[i] = (g->x_hi[i] - g->x_lo[i]) / g->grid_dims[i];
  }
  finish_dims(g, accuracy);
  return 0;
}

// createGridSetDx (src/utils/lsm_grid.c:51-112): the same spacing on every axis,
// the grid dimensions rounded from (x_hi - x_lo)/dx and x_hi moved to match.
LSMHIP_EXPORT int lsmhip_grid_set_dx(LsmhipGrid* g, int num_dims, real_t dx, const real_t* x_lo, const real_t* x_hi,
                                     int accuracy) {
  if (!check_grid_args(num_dims, accuracy, "lsmhip_grid_set_dx")) return -1;
  *g = LsmhipGrid{};
  g->num_dims = num_dims;
  for (int i = 0; i < num_dims; i++) {
    g->dx[i] = dx;
    const real_t diff = x_hi[i] - x_lo[i];
    g->grid_dims[i] = (int)std::rint(diff / dx);
    g->x_lo[i] = x_lo[i];
    g->x_hi[i] = x_lo[i] + g->grid_dims[i] * dx;
  }
  finish_dims(g, accuracy);
  return 0;
}

// writeGridToBinaryFile without compression (src/utils/lsm_grid.c:614-678).
LSMHIP_EXPORT int lsmhip_write_grid(const LsmhipGrid* g, const char* file_name) {
  File f(fopen(file_name, "wb"));
  if (!f.f) {
    set_error(kErrArg, "lsmhip_write_grid: cannot open %s for writing", file_name);
    return -1;
  }
  FILE* p = f.f;
  const int boxes[30] = {g->ilo_gb, g->ihi_gb, g->jlo_gb, g->jhi_gb, g->klo_gb, g->khi_gb,
                         g->ilo_fb, g->ihi_fb, g->jlo_fb, g->jhi_fb, g->klo_fb, g->khi_fb,
                         g->ilo_D1_fb, g->ihi_D1_fb, g->jlo_D1_fb, g->jhi_D1_fb, g->klo_D1_fb, g->khi_D1_fb,
                         g->ilo_D2_fb, g->ihi_D2_fb, g->jlo_D2_fb, g->jhi_D2_fb, g->klo_D2_fb, g->khi_D2_fb,
                         g->ilo_D3_fb, g->ihi_D3_fb, g->jlo_D3_fb, g->jhi_D3_fb, g->klo_D3_fb, g->khi_D3_fb};
  const unsigned char marks[5] = {g->mark_gb, g->mark_D1, g->mark_D2, g->mark_D3, g->mark_fb};
  const bool ok = put(p, &g->num_dims, 1) && put(p, g->x_lo, 3) && put(p, g->x_hi, 3) && put(p, g->x_lo_ghostbox, 3) &&
                  put(p, g->x_hi_ghostbox, 3) && put(p, g->grid_dims, 3) && put(p, g->grid_dims_ghostbox, 3) &&
                  put(p, g->dx, 3) && put(p, &g->num_gridpts, 1) && put(p, boxes, 30) &&
                  put(p, &g->num_nb_levels, 1) && put(p, marks, 5) && put(p, &g->beta, 1) && put(p, &g->gamma, 1);
  if (!ok) {
    set_error(kErrArg, "lsmhip_write_grid: short write to %s", file_name);
    return -1;
  }
  return 0;
}

// readGridFromBinaryFile without compression (src/utils/lsm_grid.c:680-763).
LSMHIP_EXPORT int lsmhip_read_grid(LsmhipGrid* g, const char* file_name) {
  File f(fopen(file_name, "rb"));
  if (!f.f) {
    set_error(kErrArg, "lsmhip_read_grid: cannot open %s", file_name);
    return -1;
  }
  FILE* p = f.f;
  int boxes[30];
  unsigned char marks[5];
  *g = LsmhipGrid{};
  const bool ok = get(p, &g->num_dims, 1) && get(p, g->x_lo, 3) && get(p, g->x_hi, 3) && get(p, g->x_lo_ghostbox, 3) &&
                  get(p, g->x_hi_ghostbox, 3) && get(p, g->grid_dims, 3) && get(p, g->grid_dims_ghostbox, 3) &&
                  get(p, g->dx, 3) && get(p, &g->num_gridpts, 1) && get(p, boxes, 30) &&
                  get(p, &g->num_nb_levels, 1) && get(p, marks, 5) && get(p, &g->beta, 1) && get(p, &g->gamma, 1);
  if (!ok) {
    set_error(kErrArg, "lsmhip_read_grid: %s is shorter than a grid record", file_name);
    return -1;
  }
  int* dst[30] = {&g->ilo_gb, &g->ihi_gb, &g->jlo_gb, &g->jhi_gb, &g->klo_gb, &g->khi_gb,
                  &g->ilo_fb, &g->ihi_fb, &g->jlo_fb, &g->jhi_fb, &g->klo_fb, &g->khi_fb,
                  &g->ilo_D1_fb, &g->ihi_D1_fb, &g->jlo_D1_fb, &g->jhi_D1_fb, &g->klo_D1_fb, &g->khi_D1_fb,
                  &g->ilo_D2_fb, &g->ihi_D2_fb, &g->jlo_D2_fb, &g->jhi_D2_fb, &g->klo_D2_fb, &g->khi_D2_fb,
                  &g->ilo_D3_fb, &g->ihi_D3_fb, &g->jlo_D3_fb, &g->jhi_D3_fb, &g->klo_D3_fb, &g->khi_D3_fb};
  for (int q = 0; q < 30; q++) *dst[q] = boxes[q];
  g->mark_gb = marks[0]; g->mark_D1 = marks[1]; g->mark_D2 = marks[2]; g->mark_D3 = marks[3]; g->mark_fb = marks[4];
  return 0;
}
