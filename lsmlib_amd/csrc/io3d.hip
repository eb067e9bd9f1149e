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
//
// Two sets of entry points: the lsmhip_* forms (a caller's Grid struct, an
// explicit on_device flag, no compression), and the reference's own names
// (createGridSetDx, writeDataArray, readGridFromBinaryFile, ...) with its
// signatures, calloc'd grids, malloc'd arrays and zip_status compression,
// which runs gzip / bzip2 on the file as src/utils/lsm_file.c:19-70 does
// (here as a child process with an argument vector, not through a shell).
#include <spawn.h>
#include <sys/wait.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

extern char** environ;

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

// src/utils/lsm_grid.c:60-70, 205-215: a dimension other than 2 or 3 becomes 2
int reference_num_dims(int num_dims) {
  if (num_dims == 2 || num_dims == 3) return num_dims;
  printf("\nInvalid grid dimension %d, set to default 2.\n", num_dims);
  return 2;
}

// Runs `tool -f file` as a child process and waits for it.
bool run_tool(const char* tool, const char* file) {
  char* argv[] = {const_cast<char*>(tool), const_cast<char*>("-f"), const_cast<char*>(file), nullptr};
  pid_t pid;
  if (posix_spawnp(&pid, tool, nullptr, nullptr, argv, environ) != 0) return false;
  int st = 0;
  if (waitpid(pid, &st, 0) < 0) return false;
  return WIFEXITED(st) && WEXITSTATUS(st) == 0;
}

// zipFile (src/utils/lsm_file.c:57-70)
void zip_file(const char* file, int zip_status, const char* who) {
  const char* tool = zip_status == GZIP ? "gzip" : zip_status == BZIP2 ? "bzip2" : nullptr;
  if (tool && !run_tool(tool, file)) set_error(kErrArg, "%s: %s -f %s failed", who, tool, file);
}

// checkUnzipFile (src/utils/lsm_file.c:19-55): a name containing ".gz" or
// ".bz2" is decompressed in place and the name without the suffix returned.
std::string unzip_file(const char* name, int* zip_status, const char* who) {
  *zip_status = NO_ZIP;
  const char* tool = nullptr;
  size_t cut = 0;
  if (strstr(name, ".gz")) {
    *zip_status = GZIP;
    tool = "gunzip";
    cut = 3;
  } else if (strstr(name, ".bz2")) {
    *zip_status = BZIP2;
    tool = "bunzip2";
    cut = 4;
  }
  if (!tool) return std::string(name);
  if (!run_tool(tool, name)) set_error(kErrArg, "%s: %s -f %s failed", who, tool, name);
  const size_t n = strlen(name);
  return std::string(name, n >= cut ? n - cut : 0);
}

// Whether p is device memory (then copied through the host); host memory,
// pinned or not, and pointers HIP does not know are read directly.
int on_device_memory(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return a.type == hipMemoryTypeDevice ? 1 : 0;
}

bool write_array_file(const real_t* data, const int* dims, long long n, const char* file_name, int on_device,
                      const char* who) {
  File f(fopen(file_name, "wb"));
  if (!f.f) {
    set_error(kErrArg, "%s: cannot open %s for writing", who, file_name);
    return false;
  }
  if (!put(f.f, dims, 3)) {
    set_error(kErrArg, "%s: short write to %s", who, file_name);
    return false;
  }
  std::vector<real_t> buf((size_t)std::min(std::max(n, 1LL), kChunk));
  for (long long o = 0; o < n; o += kChunk) {
    const long long m = std::min(kChunk, n - o);
    if (!to_host(buf.data(), data + o, m, on_device, who)) return false;
    if (!put(f.f, buf.data(), (size_t)m)) {
      set_error(kErrArg, "%s: short write to %s", who, file_name);
      return false;
    }
  }
  return true;
}

bool write_grid_file(const LsmhipGrid* g, const char* file_name, const char* who);
bool read_grid_file(LsmhipGrid* g, FILE* p, const char* who, const char* file_name);

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
  if (!data || !grid_dims_ghostbox || !file_name) {
    set_error(kErrArg, "%s: NULL argument", who);
    return -1;
  }
  if (grid_dims_ghostbox[0] < 0 || grid_dims_ghostbox[1] < 0 || grid_dims_ghostbox[2] < 0) {
    set_error(kErrArg, "%s: negative dimension (%d, %d, %d)", who, grid_dims_ghostbox[0], grid_dims_ghostbox[1],
              grid_dims_ghostbox[2]);
    return -1;
  }
  const long long n = (long long)grid_dims_ghostbox[0] * grid_dims_ghostbox[1] * grid_dims_ghostbox[2];
  return write_array_file(data, grid_dims_ghostbox, n, file_name, on_device, who) ? 0 : -1;
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
  if (grid_dims_ghostbox[0] < 0 || grid_dims_ghostbox[1] < 0 || grid_dims_ghostbox[2] < 0) {
    set_error(kErrArg, "%s: %s has a negative dimension (%d, %d, %d)", who, file_name, grid_dims_ghostbox[0],
              grid_dims_ghostbox[1], grid_dims_ghostbox[2]);
    return -1;
  }
  const long long n = (long long)grid_dims_ghostbox[0] * grid_dims_ghostbox[1] * grid_dims_ghostbox[2];
  if (n > capacity) {
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

namespace {
// The grid record (src/utils/lsm_grid.c:614-678): every field in declaration
// order with its own size, no padding.
bool write_grid_file(const LsmhipGrid* g, const char* file_name, const char* who) {
  File f(fopen(file_name, "wb"));
  if (!f.f) {
    set_error(kErrArg, "%s: cannot open %s for writing", who, file_name);
    return false;
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
  if (!ok) set_error(kErrArg, "%s: short write to %s", who, file_name);
  return ok;
}

// readGridFromBinaryFile's record (src/utils/lsm_grid.c:680-763) into g.
bool read_grid_file(LsmhipGrid* g, FILE* p, const char* who, const char* file_name) {
  int boxes[30] = {};
  unsigned char marks[5] = {};
  const bool ok = get(p, &g->num_dims, 1) && get(p, g->x_lo, 3) && get(p, g->x_hi, 3) && get(p, g->x_lo_ghostbox, 3) &&
                  get(p, g->x_hi_ghostbox, 3) && get(p, g->grid_dims, 3) && get(p, g->grid_dims_ghostbox, 3) &&
                  get(p, g->dx, 3) && get(p, &g->num_gridpts, 1) && get(p, boxes, 30) &&
                  get(p, &g->num_nb_levels, 1) && get(p, marks, 5) && get(p, &g->beta, 1) && get(p, &g->gamma, 1);
  int* dst[30] = {&g->ilo_gb, &g->ihi_gb, &g->jlo_gb, &g->jhi_gb, &g->klo_gb, &g->khi_gb,
                  &g->ilo_fb, &g->ihi_fb, &g->jlo_fb, &g->jhi_fb, &g->klo_fb, &g->khi_fb,
                  &g->ilo_D1_fb, &g->ihi_D1_fb, &g->jlo_D1_fb, &g->jhi_D1_fb, &g->klo_D1_fb, &g->khi_D1_fb,
                  &g->ilo_D2_fb, &g->ihi_D2_fb, &g->jlo_D2_fb, &g->jhi_D2_fb, &g->klo_D2_fb, &g->khi_D2_fb,
                  &g->ilo_D3_fb, &g->ihi_D3_fb, &g->jlo_D3_fb, &g->jhi_D3_fb, &g->klo_D3_fb, &g->khi_D3_fb};
  for (int q = 0; q < 30; q++) *dst[q] = boxes[q];
  g->mark_gb = marks[0]; g->mark_D1 = marks[1]; g->mark_D2 = marks[2]; g->mark_D3 = marks[3]; g->mark_fb = marks[4];
  if (!ok) set_error(kErrArg, "%s: %s is shorter than a grid record", who, file_name);
  return ok;
}
}  // namespace

// writeGridToBinaryFile without compression (src/utils/lsm_grid.c:614-678).
LSMHIP_EXPORT int lsmhip_write_grid(const LsmhipGrid* g, const char* file_name) {
  return write_grid_file(g, file_name, "lsmhip_write_grid") ? 0 : -1;
}

// readGridFromBinaryFile without compression (src/utils/lsm_grid.c:680-763).
LSMHIP_EXPORT int lsmhip_read_grid(LsmhipGrid* g, const char* file_name) {
  File f(fopen(file_name, "rb"));
  if (!f.f) {
    set_error(kErrArg, "lsmhip_read_grid: cannot open %s", file_name);
    return -1;
  }
  *g = LsmhipGrid{};
  return read_grid_file(g, f.f, "lsmhip_read_grid", file_name) ? 0 : -1;
}

// ---------------------------------------------------------------------------
// The reference's own names and signatures (include/lsm_grid.h,
// include/lsm_data_arrays.h).
// ---------------------------------------------------------------------------
namespace {
Grid* grid_from(int num_dims, const int* grid_dims, const real_t* dxs, const real_t* x_lo, const real_t* x_hi,
                int accuracy, const char* who) {
  if (accuracy < 0 || accuracy > 3) {
    set_error(kErrArg, "%s: accuracy must be LOW (0) .. VERY_HIGH (3), got %d", who, accuracy);
    return nullptr;
  }
  if (!x_lo || !x_hi || (!grid_dims && !dxs)) {
    set_error(kErrArg, "%s: NULL argument", who);
    return nullptr;
  }
  Grid* g = (Grid*)calloc(1, sizeof(Grid));  // allocateGrid, src/utils/lsm_grid.c:35-39
  if (!g) {
    set_error(kErrArg, "%s: out of host memory", who);
    return nullptr;
  }
  g->num_dims = reference_num_dims(num_dims);
  for (int i = 0; i < g->num_dims; i++) {
    g->x_lo[i] = x_lo[i];
    if (grid_dims) {  // createGridSetGridDims, src/utils/lsm_grid.c:189-232
      g->grid_dims[i] = grid_dims[i];
      g->x_hi[i] = x_hi[i];
      g->dx[i] = (g->x_hi[i] - g->x_lo[i]) / g->grid_dims[i];
    } else {  // createGridSetDx / SetDxDyDz, src/utils/lsm_grid.c:51-187
      g->dx[i] = dxs[i];
      const real_t diff = x_hi[i] - x_lo[i];
      g->grid_dims[i] = (int)std::rint(diff / g->dx[i]);
      g->x_hi[i] = x_lo[i] + g->grid_dims[i] * g->dx[i];
    }
  }
  finish_dims(g, accuracy);
  return g;
}
}  // namespace

LSMHIP_EXPORT Grid* createGridSetDx(int num_dims, real_t dx, real_t* x_lo, real_t* x_hi,
                                    LSMLIB_SPATIAL_DERIVATIVE_ACCURACY_TYPE accuracy) {
  const real_t d[3] = {dx, dx, dx};
  return grid_from(num_dims, nullptr, d, x_lo, x_hi, (int)accuracy, "createGridSetDx");
}

LSMHIP_EXPORT Grid* createGridSetDxDyDz(int num_dims, real_t dx, real_t dy, real_t dz, real_t* x_lo, real_t* x_hi,
                                        LSMLIB_SPATIAL_DERIVATIVE_ACCURACY_TYPE accuracy) {
  const real_t d[3] = {dx, dy, dz};
  return grid_from(num_dims, nullptr, d, x_lo, x_hi, (int)accuracy, "createGridSetDxDyDz");
}

LSMHIP_EXPORT Grid* createGridSetGridDims(int num_dims, int* grid_dims, real_t* x_lo, real_t* x_hi,
                                          LSMLIB_SPATIAL_DERIVATIVE_ACCURACY_TYPE accuracy) {
  if (!grid_dims) {
    set_error(kErrArg, "createGridSetGridDims: NULL grid_dims");
    return nullptr;
  }
  return grid_from(num_dims, grid_dims, nullptr, x_lo, x_hi, (int)accuracy, "createGridSetGridDims");
}

LSMHIP_EXPORT void destroyGrid(Grid* grid) { free(grid); }

LSMHIP_EXPORT void writeGridToBinaryFile(Grid* grid, char* file_name, int zip_status) {
  const char* who = "writeGridToBinaryFile";
  if (!grid || !file_name) {
    set_error(kErrArg, "%s: NULL argument", who);
    return;
  }
  if (write_grid_file(grid, file_name, who)) zip_file(file_name, zip_status, who);
}

LSMHIP_EXPORT Grid* readGridFromBinaryFile(char* file_name) {
  const char* who = "readGridFromBinaryFile";
  if (!file_name) {
    set_error(kErrArg, "%s: NULL file name", who);
    return nullptr;
  }
  int zip = NO_ZIP;
  const std::string base = unzip_file(file_name, &zip, who);
  Grid* g = (Grid*)calloc(1, sizeof(Grid));
  if (!g) {
    set_error(kErrArg, "%s: out of host memory", who);
    return nullptr;
  }
  FILE* p = fopen(base.c_str(), "rb");
  if (!p) {  // the reference returns the zeroed grid
    printf("\nCould not open file %s", base.c_str());
    set_error(kErrArg, "%s: cannot open %s", who, base.c_str());
    return g;
  }
  read_grid_file(g, p, who, base.c_str());
  fclose(p);
  zip_file(base.c_str(), zip, who);
  return g;
}

LSMHIP_EXPORT void writeDataArray(real_t* data, Grid* grid, char* file_name, int zip_status) {
  const char* who = "writeDataArray";
  if (!data || !grid || !file_name) {
    set_error(kErrArg, "%s: NULL argument", who);
    return;
  }
  if (grid->num_gridpts < 0) {
    set_error(kErrArg, "%s: grid holds %d points", who, grid->num_gridpts);
    return;
  }
  if (write_array_file(data, grid->grid_dims_ghostbox, grid->num_gridpts, file_name, on_device_memory(data), who))
    zip_file(file_name, zip_status, who);
}

LSMHIP_EXPORT void writeDataArrayNoGrid(real_t* data, int* n, char* file_name, int zip_status) {
  const char* who = "writeDataArrayNoGrid";
  if (!data || !n || !file_name) {
    set_error(kErrArg, "%s: NULL argument", who);
    return;
  }
  if (n[0] < 0 || n[1] < 0 || n[2] < 0) {
    set_error(kErrArg, "%s: negative dimension (%d, %d, %d)", who, n[0], n[1], n[2]);
    return;
  }
  const long long m = (long long)n[0] * n[1] * n[2];
  if (write_array_file(data, n, m, file_name, on_device_memory(data), who)) zip_file(file_name, zip_status, who);
}

LSMHIP_EXPORT real_t* readDataArray(int* grid_dims, char* file_name) {
  const char* who = "readDataArray";
  if (!grid_dims || !file_name) {
    set_error(kErrArg, "%s: NULL argument", who);
    return nullptr;
  }
  int zip = NO_ZIP;
  const std::string base = unzip_file(file_name, &zip, who);
  FILE* p = fopen(base.c_str(), "rb");
  if (!p) {
    printf("\nCould not open file %s", file_name);
    set_error(kErrArg, "%s: cannot open %s", who, file_name);
    return nullptr;
  }
  real_t* data = nullptr;
  if (!get(p, grid_dims, 3)) {
    set_error(kErrArg, "%s: cannot read the header of %s", who, base.c_str());
  } else if (grid_dims[0] < 0 || grid_dims[1] < 0 || grid_dims[2] < 0) {
    set_error(kErrArg, "%s: %s has a negative dimension (%d, %d, %d)", who, base.c_str(), grid_dims[0],
              grid_dims[1], grid_dims[2]);
  } else {
    const long long n = (long long)grid_dims[0] * grid_dims[1] * grid_dims[2];
    data = (real_t*)malloc(sizeof(real_t) * (size_t)(n > 0 ? n : 1));
    if (!data) {
      set_error(kErrArg, "%s: out of host memory for %lld reals", who, n);
    } else if (!get(p, data, (size_t)n)) {
      set_error(kErrArg, "%s: %s ends before its %lld reals", who, base.c_str(), n);
      free(data);
      data = nullptr;
    }
  }
  fclose(p);
  zip_file(base.c_str(), zip, who);
  return data;
}
