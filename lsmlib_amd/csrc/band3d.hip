// band3d.hip -- narrow-band construction and the band checks around it
// (src/toolbox/lsm_localization3d.f:48-653, and the 2D twins of
// src/toolbox/lsm_localization2d.f:48-772), with index lists bit-exact to the
// reference's sequential loops.  A 2D band is a 3D band of one plane (k = 0)
// with no index_z list and four neighbour directions.
//
// Level 0 (lsm3dDetermineNarrowBand, :174-297) is the set |phi| < width in
// (k,j,i)-lexicographic order of phi's ghost box, i.e. memory order: an
// ordered stream compaction.  One pass counts the points of each 4096-cell
// chunk (band, outer-layer phi <= 0, outer-layer phi > 0), one block scans the
// chunk counts, and a second pass recomputes the flags and scatters each point
// to its rank.  index_outer receives minus points from the front in order and
// plus points from the back in reverse order, as the reference's two counters do.
//
// Levels 1..L (lsm3dMarkNarrowBandNeighbors, :48-167) are grown by first claim:
// the reference visits parents m of level L-1 in list order and, for each, the
// neighbours +x,-x,+y,-y,+z,-z, appending every still-unmarked neighbour.  A
// neighbour therefore belongs to the smallest candidate c = 6 (m - n_lo) + dir
// that reaches it.  Here every candidate claims its target with atomicMin on a
// per-cell claim word; the winning candidates are then compacted in candidate
// order, which reproduces the reference's list order exactly.
#include "lsmhip_core.hpp"

using namespace lsmhip;

namespace {

constexpr int kT = 256;
constexpr int kPer = 16;               // items per thread in a chunk
static_assert(3 * kPer <= 64, "level-0 flags of a thread are kept in one 64-bit mask");
constexpr int kChunk = kT * kPer;      // items per block
constexpr int kNoClaim = 0x7fffffff;

__device__ __forceinline__ int wave_prefix(bool f, int* total) {
  const unsigned long long m = __ballot(f);
  *total = __popcll(m);
  return __popcll(m & ((1ull << (threadIdx.x & 63)) - 1ull));
}

// ----- level 0 ------------------------------------------------------------------------------------
struct Level0 {
  const real_t* phi;  // over box B (phi's ghost box), memory order
  const real_t* mask; // from-mask variant: mask over B
  long long n;        // cells of B
  int nx, ny;
  Box B;
  real_t width, width_inner;
  int mode;           // 0: |phi| < width; 1: mask >= 0; 2: mask <= 0; 3: mask < 0
  // the value that decides cell t (phi, or the mask), 0 past the end
  __device__ real_t load(long long t) const {
    if (t >= n) return real_t(0);
    return mode == 0 ? phi[t] : mask[t];
  }
  __device__ void flags_of(long long t, real_t v, bool f[3]) const {
    f[0] = f[1] = f[2] = false;
    if (t >= n) return;
    if (mode == 0) {
      const real_t a = fabs(v);
      f[0] = a < width;
      const bool outer = f[0] && a >= width_inner;
      f[1] = outer && v <= real_t(0);
      f[2] = outer && !(v <= real_t(0));
    } else {
      f[0] = mode == 1 ? v >= real_t(0) : (mode == 2 ? v <= real_t(0) : v < real_t(0));
    }
  }
};

// Per-block placement of up to three flags over kPer passes of kT items
// (item of pass it and thread x: base + it * kT + x).  Pass 1 loads and flags
// every item, and each wave's count per pass and flag goes to LDS; after one
// barrier, pass 2 places the items of each pass in item order: the block
// offset is the counts of the earlier passes and of the earlier waves of the
// same pass, plus the lane's prefix within its wave (a ballot).
struct PassCounts {
  int c[kPer][kT / 64][3];
};
__device__ __forceinline__ void record_pass(PassCounts& pc, int it, const bool f[3]) {
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int tot = __popcll(__ballot(f[q]));
    if ((threadIdx.x & 63) == 0) pc.c[it][w][q] = tot;
  }
}
// Block offsets of this thread's items of pass it (flags f) given the totals
// of the passes before it (run, advanced past pass it on return).  Every lane
// of the wave must call it (it ballots).
__device__ __forceinline__ void place_pass(const PassCounts& pc, int it, const bool f[3], int run[3], int off[3]) {
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    int tot;
    const int lane_pre = wave_prefix(f[q], &tot);
    int before = 0, all = 0;
#pragma unroll
    for (int v = 0; v < kT / 64; v++) {
      const int x = pc.c[it][v][q];
      before += v < w ? x : 0;
      all += x;
    }
    off[q] = run[q] + before + lane_pre;
    run[q] += all;
  }
}

__global__ void __launch_bounds__(kT) level0_count_kernel(Level0 a, int* counts) {
  const long long base = (long long)blockIdx.x * kChunk;
  int c[3] = {0, 0, 0};
  real_t v[kPer];  // every load of the chunk issued before any flag is formed
#pragma unroll
  for (int it = 0; it < kPer; it++) v[it] = a.load(base + it * kT + threadIdx.x);
#pragma unroll
  for (int it = 0; it < kPer; it++) {
    bool f[3];
    a.flags_of(base + it * kT + threadIdx.x, v[it], f);
    for (int q = 0; q < 3; q++) c[q] += f[q];
  }
  for (int q = 0; q < 3; q++)
    for (int o = 32; o > 0; o >>= 1) c[q] += __shfl_xor(c[q], o, 64);
  __shared__ int s[3][kT / 64];
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < 3; q++) s[q][threadIdx.x >> 6] = c[q];
  __syncthreads();
  if (threadIdx.x < 3) {
    int t = 0;
    for (int v = 0; v < kT / 64; v++) t += s[threadIdx.x][v];
    counts[3 * blockIdx.x + threadIdx.x] = t;
  }
}

// Exclusive scan of n interleaved triples (one block); totals to tot[0..2].
// Each thread sums a contiguous segment, the 1024 segment sums are scanned
// in parallel (shuffles inside each wave, then the 16 wave totals), and each
// thread rewrites its segment from its offset.
__global__ void __launch_bounds__(1024) scan3_kernel(int* v, long long n, long long* tot) {
  __shared__ long long wsum[3][1024 / 64];
  const long long per = (n + 1023) / 1024;
  const long long lo = threadIdx.x * per, hi = min(n, lo + per);
  long long s[3] = {0, 0, 0};
  for (long long t = lo; t < hi; t++)
    for (int q = 0; q < 3; q++) s[q] += v[3 * t + q];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  long long inc[3];
  for (int q = 0; q < 3; q++) {
    long long x = s[q];
    for (int o = 1; o < 64; o <<= 1) {
      const long long y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    inc[q] = x;  // inclusive within the wave
  }
  if (lane == 63)
    for (int q = 0; q < 3; q++) wsum[q][w] = inc[q];
  __syncthreads();
  long long run[3];
  for (int q = 0; q < 3; q++) {
    long long before = 0, all = 0;
    for (int u = 0; u < 1024 / 64; u++) {
      if (u < w) before += wsum[q][u];
      all += wsum[q][u];
    }
    run[q] = before + inc[q] - s[q];
    if (threadIdx.x == 0) tot[q] = all;
  }
  for (long long t = lo; t < hi; t++)
    for (int q = 0; q < 3; q++) {
      const int x = v[3 * t + q];
      v[3 * t + q] = (int)run[q];
      run[q] += x;
    }
}

struct Level0Out {
  int *ix, *iy, *iz;  // element 0 is index nlo; iz null for a 2D band
  int nlo;
  int* outer;         // element 0 is index nlo_outer
  int nlo_outer, nhi_outer;
  unsigned char* nb;
  Box nbb;
};

__global__ void __launch_bounds__(kT) level0_write_kernel(Level0 a, const int* offsets, Level0Out o) {
  __shared__ PassCounts pc;
  const long long base = (long long)blockIdx.x * kChunk;
  const long long t0 = base + threadIdx.x;
  unsigned long long fl = 0;  // bit 3 it + q: flag q of pass it (48 bits)
  {
    real_t v[kPer];  // every load of the chunk issued before any flag is formed
#pragma unroll
    for (int it = 0; it < kPer; it++) v[it] = a.load(t0 + it * kT);
#pragma unroll
    for (int it = 0; it < kPer; it++) {
      bool f[3];
      a.flags_of(t0 + it * kT, v[it], f);
      record_pass(pc, it, f);
      for (int q = 0; q < 3; q++) fl |= (unsigned long long)f[q] << (3 * it + q);
    }
  }
  __syncthreads();
  int run[3] = {offsets[3 * blockIdx.x], offsets[3 * blockIdx.x + 1], offsets[3 * blockIdx.x + 2]};
  // (i, j, k) of the thread's first cell, then advanced by kT cells per pass:
  // one division per thread instead of three per cell
  int i = a.B.ilo + (int)(t0 % a.nx);
  int j = a.B.jlo + (int)((t0 / a.nx) % a.ny);
  int k = a.B.klo + (int)(t0 / ((long long)a.nx * a.ny));
  const bool same = o.nbb.ilo == a.B.ilo && o.nbb.ihi == a.B.ihi && o.nbb.jlo == a.B.jlo && o.nbb.jhi == a.B.jhi &&
                    o.nbb.klo == a.B.klo && o.nbb.khi == a.B.khi;
  for (int it = 0; it < kPer; it++) {
    const long long t = t0 + it * kT;
    const bool f[3] = {((fl >> (3 * it)) & 1ull) != 0, ((fl >> (3 * it + 1)) & 1ull) != 0,
                       ((fl >> (3 * it + 2)) & 1ull) != 0};
    int off[3];
    place_pass(pc, it, f, run, off);
    if (t < a.n) {
      const long long nbo = same ? t
                                 : (long long)(i - o.nbb.ilo) +
                                       (long long)o.nbb.nx() * ((j - o.nbb.jlo) + (long long)o.nbb.ny() * (k - o.nbb.klo));
      o.nb[nbo] = f[0] ? 1 : 0;
      if (f[0]) {
        const int r = off[0];
        o.ix[r] = i;
        o.iy[r] = j;
        if (o.iz) o.iz[r] = k;
        if (f[1]) o.outer[off[1]] = o.nlo + r;
        if (f[2]) o.outer[(o.nhi_outer - o.nlo_outer) - off[2]] = o.nlo + r;
      }
    }
    i += kT;
    while (i > a.B.ihi) {
      i -= a.nx;
      if (++j > a.B.jhi) {
        j = a.B.jlo;
        ++k;
      }
    }
  }
}

// ----- levels >= 1 ------------------------------------------------------------------------------
struct Grow {
  const int *ix, *iy, *iz;  // element 0 is index base; iz null for a 2D band
  int base;
  int plo;                  // first parent index
  long long ncand;          // directions (6 in 3D, 4 in 2D) x parents
  unsigned char* nb;
  Box nbb;
  int* claim;               // over nbb
  unsigned char* win;       // per candidate: tried to claim a free cell (claim pass), then won it (count pass)
};

// NDIR (6 in 3D, 4 in 2D) is a template argument, so the candidate's split
// into (parent, direction) is a division by a constant.
template <int NDIR>
__device__ __forceinline__ bool target(const Grow& g, long long c, int& off, int& qi, int& qj, int& qk, int* flag) {
  if (c >= g.ncand) return false;
  const int m = g.plo + (int)(c / NDIR), d = (int)(c % NDIR);
  const int i = g.ix[m - g.base], j = g.iy[m - g.base], k = g.iz ? g.iz[m - g.base] : 0;
  const Box& b = g.nbb;
  if (!b.has(i, j, k)) {
    *flag = 1;
    return false;
  }
  qi = i;
  qj = j;
  qk = k;
  switch (d) {
    case 0: if (!(i < b.ihi)) return false; qi = i + 1; break;
    case 1: if (!(i > b.ilo)) return false; qi = i - 1; break;
    case 2: if (!(j < b.jhi)) return false; qj = j + 1; break;
    case 3: if (!(j > b.jlo)) return false; qj = j - 1; break;
    case 4: if (!(k < b.khi)) return false; qk = k + 1; break;
    default: if (!(k > b.klo)) return false; qk = k - 1; break;
  }
  off = (qi - b.ilo) + b.nx() * ((qj - b.jlo) + b.ny() * (qk - b.klo));
  return true;
}

// One thread per parent: its index read once, the marks of its NDIR
// neighbours loaded together, then a claim for each free one (candidate
// NDIR * parent + direction, as target() numbers them).
template <int NDIR>
__global__ void __launch_bounds__(kT) claim_kernel(Grow g, int* flag) {
  const long long p = (long long)blockIdx.x * kT + threadIdx.x;
  if (p >= g.ncand / NDIR) return;
  const long long c0 = NDIR * p;
  const int m = g.plo + (int)p;
  const int i = g.ix[m - g.base], j = g.iy[m - g.base], k = g.iz ? g.iz[m - g.base] : 0;
  const Box& b = g.nbb;
  if (!b.has(i, j, k)) {
    *flag = 1;
#pragma unroll
    for (int d = 0; d < NDIR; d++) g.win[c0 + d] = 0;
    return;
  }
  const bool ok[6] = {i < b.ihi, i > b.ilo, j < b.jhi, j > b.jlo, k < b.khi, k > b.klo};
  const int sy = b.nx(), sz = b.nx() * b.ny();
  const int o = (i - b.ilo) + sy * ((j - b.jlo) + b.ny() * (k - b.klo));
  const int step[6] = {1, -1, sy, -sy, sz, -sz};
  unsigned char mk[NDIR];
#pragma unroll
  for (int d = 0; d < NDIR; d++) mk[d] = ok[d] ? g.nb[o + step[d]] : 1;
#pragma unroll
  for (int d = 0; d < NDIR; d++) {
    const bool tried = ok[d] && mk[d] == 0;
    if (tried) atomicMin(&g.claim[o + step[d]], (int)(c0 + d));
    g.win[c0 + d] = tried ? 1 : 0;  // the count pass looks only at these
  }
}

template <int NDIR>
__global__ void __launch_bounds__(kT) grow_count_kernel(Grow g, int* counts, int* flag) {
  const long long base = (long long)blockIdx.x * kChunk;
  int c = 0;
  for (int it = 0; it < kPer; it++) {
    const long long cand = base + it * kT + threadIdx.x;
    // only candidates that tried to claim a free cell can have won it
    bool w = false;
    if (cand < g.ncand && g.win[cand]) {
      int off, qi, qj, qk;
      w = target<NDIR>(g, cand, off, qi, qj, qk, flag) && g.claim[off] == (int)cand;
      g.win[cand] = w ? 1 : 0;
    }
    c += w;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  __shared__ int s[kT / 64];
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int v = 0; v < kT / 64; v++) t += s[v];
    counts[3 * blockIdx.x] = t;
    counts[3 * blockIdx.x + 1] = 0;
    counts[3 * blockIdx.x + 2] = 0;
  }
}

template <int NDIR>
__global__ void __launch_bounds__(kT) grow_write_kernel(Grow g, const int* offsets, int* ox, int* oy, int* oz,
                                                        int start_rel, unsigned char mark, int* flag) {
  __shared__ PassCounts pc;
  const long long base = (long long)blockIdx.x * kChunk;
  const long long c0 = base + threadIdx.x;
  unsigned wins = 0;  // bit it: the count pass's result for pass it, all loaded first
  {
    unsigned char wb[kPer];
#pragma unroll
    for (int it = 0; it < kPer; it++) wb[it] = c0 + it * kT < g.ncand ? g.win[c0 + it * kT] : 0;
#pragma unroll
    for (int it = 0; it < kPer; it++) {
      const bool f[3] = {wb[it] != 0, false, false};
      record_pass(pc, it, f);
      wins |= (unsigned)(wb[it] != 0) << it;
    }
  }
  __syncthreads();
  int run[3] = {offsets[3 * blockIdx.x], 0, 0};
  for (int it = 0; it < kPer; it++) {
    const bool f[3] = {((wins >> it) & 1u) != 0, false, false};
    int off[3];
    place_pass(pc, it, f, run, off);
    if (!f[0]) continue;
    const long long cand = c0 + it * kT;
    int off_cell = 0, qi = 0, qj = 0, qk = 0;
    if (!target<NDIR>(g, cand, off_cell, qi, qj, qk, flag)) continue;  // cannot happen for a winner
    const int r = start_rel + off[0];
    ox[r] = qi;
    oy[r] = qj;
    if (oz) oz[r] = qk;
    g.nb[off_cell] = mark;
    g.claim[off_cell] = kNoClaim;
  }
}

__global__ void __launch_bounds__(256) fill_int_kernel(int* p, long long n, int v) {
  for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long long)gridDim.x * 256) p[t] = v;
}

// ----- device scratch of the band builder ---------------------------------------------------------
int* g_claim = nullptr;
long long g_claim_n = 0;
unsigned char* g_win = nullptr;
long long g_win_n = 0;
int* g_counts = nullptr;
long long g_counts_n = 0;
long long* g_tot = nullptr;

bool ensure_win(const char* who, long long n) {
  if (n <= g_win_n) return true;
  if (g_win) (void)hipFree(g_win);
  g_win = nullptr;
  g_win_n = 0;
  if (!check_hip(hipMalloc(&g_win, (size_t)n), who)) return false;
  g_win_n = n;
  return true;
}

bool ensure(const char* who, long long claim_n, long long counts_n) {
  if (claim_n > g_claim_n) {
    if (g_claim) (void)hipFree(g_claim);
    g_claim = nullptr;
    g_claim_n = 0;
    if (!check_hip(hipMalloc(&g_claim, sizeof(int) * claim_n), who)) return false;
    g_claim_n = claim_n;
  }
  if (counts_n > g_counts_n) {
    if (g_counts) (void)hipFree(g_counts);
    g_counts = nullptr;
    g_counts_n = 0;
    if (!check_hip(hipMalloc(&g_counts, sizeof(int) * 3 * counts_n), who)) return false;
    g_counts_n = counts_n;
  }
  if (!g_tot && !check_hip(hipMalloc(&g_tot, sizeof(long long) * 4), who)) return false;
  return true;
}

bool read_totals(const char* who, long long tot[3]) {
  if (!check_hip(hipMemcpyAsync(tot, g_tot, sizeof(long long) * 3, hipMemcpyDeviceToHost, stream()), who)) return false;
  return check_hip(hipStreamSynchronize(stream()), who);
}

// Grows levels 1..level from level 0 (n_lo/n_hi are host arrays of the caller).
// iz null: a 2D band, grown along +x, -x, +y, -y.
bool grow_levels(const char* who, unsigned char* nb, const Box& nbb, int* ix, int* iy, int* iz, int nlo, int nhi,
                 int* n_lo, int* n_hi, int level) {
  const int ndir = iz ? 6 : 4;
  if (level < 1) return true;
  const long long ncells = nbb.volume();
  long long maxcand = 0;
  if (!ensure(who, ncells, 1)) return false;
  int* flag = device_flag();
  if (!flag) return false;
  hipLaunchKernelGGL(fill_int_kernel, dim3(4096), dim3(256), 0, stream(), g_claim, ncells, kNoClaim);
  int next = n_hi[0] + 1;  // first free slot (the reference's running count)
  for (int L = 1; L <= level; L++) {
    if (n_lo[L - 1] < 0 || n_hi[L - 1] < n_lo[L - 1]) {  // no parents: the level stays empty
      n_lo[L] = -1;
      n_hi[L] = -1;
      continue;
    }
    const long long parents = (long long)n_hi[L - 1] - n_lo[L - 1] + 1;
    const long long ncand = ndir * parents;
    const long long nblk = (ncand + kChunk - 1) / kChunk;
    if (ncand > maxcand) maxcand = ncand;
    if (!ensure(who, ncells, nblk) || !ensure_win(who, ncand)) return false;
    Grow g{ix, iy, iz, nlo, n_lo[L - 1], ncand, nb, nbb, g_claim, g_win};
    const dim3 cgrid((unsigned)((parents + kT - 1) / kT));  // one thread per parent
    if (ndir == 6) {
      hipLaunchKernelGGL(claim_kernel<6>, cgrid, dim3(kT), 0, stream(), g, flag);
      hipLaunchKernelGGL(grow_count_kernel<6>, dim3((unsigned)nblk), dim3(kT), 0, stream(), g, g_counts, flag);
    } else {
      hipLaunchKernelGGL(claim_kernel<4>, cgrid, dim3(kT), 0, stream(), g, flag);
      hipLaunchKernelGGL(grow_count_kernel<4>, dim3((unsigned)nblk), dim3(kT), 0, stream(), g, g_counts, flag);
    }
    hipLaunchKernelGGL(scan3_kernel, dim3(1), dim3(1024), 0, stream(), g_counts, nblk, g_tot);
    long long tot[3];
    if (!check_hip(hipGetLastError(), who) || !read_totals(who, tot)) return false;
    if (tot[0] == 0) {
      n_lo[L] = -1;
      n_hi[L] = -1;
      continue;
    }
    if ((long long)next + tot[0] - 1 > nhi) {
      set_error(kErrArg, "%s: level %d needs index entries up to %lld but the lists end at %d", who, L,
                (long long)next + tot[0] - 1, nhi);
      return false;
    }
    if (ndir == 6)
      hipLaunchKernelGGL(grow_write_kernel<6>, dim3((unsigned)nblk), dim3(kT), 0, stream(), g, g_counts, ix, iy, iz,
                         next - nlo, (unsigned char)(L + 1), flag);
    else
      hipLaunchKernelGGL(grow_write_kernel<4>, dim3((unsigned)nblk), dim3(kT), 0, stream(), g, g_counts, ix, iy, iz,
                         next - nlo, (unsigned char)(L + 1), flag);
    if (!check_hip(hipGetLastError(), who)) return false;
    n_lo[L] = next;
    n_hi[L] = next + (int)tot[0] - 1;
    next += (int)tot[0];
  }
  return true;
}

bool level0(const char* who, const Level0& a, const Level0Out& o, long long tot[3]) {
  const long long nblk = (a.n + kChunk - 1) / kChunk;
  if (!ensure(who, 0, nblk)) return false;
  hipLaunchKernelGGL(level0_count_kernel, dim3((unsigned)nblk), dim3(kT), 0, stream(), a, g_counts);
  hipLaunchKernelGGL(scan3_kernel, dim3(1), dim3(1024), 0, stream(), g_counts, nblk, g_tot);
  if (!check_hip(hipGetLastError(), who) || !read_totals(who, tot)) return false;
  return true;
}

bool level0_write(const char* who, const Level0& a, const Level0Out& o) {
  const long long nblk = (a.n + kChunk - 1) / kChunk;
  hipLaunchKernelGGL(level0_write_kernel, dim3((unsigned)nblk), dim3(kT), 0, stream(), a, g_counts, o);
  return check_hip(hipGetLastError(), who);
}

}  // namespace

// ---------------------------------------------------------------------------
// src/toolbox/lsm_localization3d.h entry points.  phi, narrow_band and the
// index lists are device arrays; n_lo / n_hi and the scalar outputs are host
// memory, as the counts the caller loops over.
// ---------------------------------------------------------------------------
LSMHIP_EXPORT void lsm3ddeterminenarrowband_(const real_t* phi, LSMHIP_BOXP(g), unsigned char* nb, LSMHIP_BOXP(nbb),
                                             int* ix, int* iy, int* iz, const int* nlo, const int* nhi, int* n_lo,
                                             int* n_hi, int* outer, const int* nlo_outer, const int* nhi_outer,
                                             int* nlo_outer_plus, int* nhi_outer_plus, int* nlo_outer_minus,
                                             int* nhi_outer_minus, const real_t* width, const real_t* width_inner,
                                             const int* level) {
  LSMHIP_REQUIRE_ARRAYS(phi, nb, ix, iy, iz, outer);
  const char* who = "lsm3ddeterminenarrowband_";
  const Box B = LSMHIP_BOX(g), nbb = LSMHIP_BOX(nbb);
  if (!check_box(nbb, B, who, "narrow_band") || !check_size(B, who) || !check_size(nbb, who)) return;
  n_lo[0] = *nlo;
  *nlo_outer_minus = *nlo_outer;
  *nhi_outer_plus = *nhi_outer;
  Level0 a{phi, nullptr, B.volume(), B.nx(), B.ny(), B, *width, *width_inner, 0};
  Level0Out o{ix, iy, iz, *nlo, outer, *nlo_outer, *nhi_outer, nb, nbb};
  long long tot[3];
  if (B.empty()) {
    tot[0] = tot[1] = tot[2] = 0;
  } else {
    if (!level0(who, a, o, tot)) return;
    if (tot[0] > (long long)*nhi - *nlo + 1 || tot[1] + tot[2] > (long long)*nhi_outer - *nlo_outer + 1) {
      set_error(kErrArg, "%s: %lld band points (%lld outer) exceed the index arrays", who, tot[0], tot[1] + tot[2]);
      return;
    }
    if (!level0_write(who, a, o)) return;
  }
  if (tot[0] > 0) {
    n_hi[0] = *nlo + (int)tot[0] - 1;
    *nhi_outer_minus = *nlo_outer + (int)tot[1] - 1;
    *nlo_outer_plus = *nhi_outer - (int)tot[2] + 1;
    grow_levels(who, nb, nbb, ix, iy, iz, *nlo, *nhi, n_lo, n_hi, *level);
  } else {
    // empty band: the reference stores the unchanged counters
    n_hi[0] = *nlo;
    *nhi_outer_minus = *nlo_outer;
    *nlo_outer_plus = *nhi_outer;
  }
  finish(who);
  take_device_flag(who);
}

LSMHIP_EXPORT void lsm3ddeterminenarrowbandfrommask_(const real_t* mask, LSMHIP_BOXP(mg), unsigned char* nb,
                                                     LSMHIP_BOXP(nbb), int* ix, int* iy, int* iz, const int* nlo,
                                                     const int* nhi, int* n_lo, int* n_hi, const int* level,
                                                     const int* use_mask_sign) {
  LSMHIP_REQUIRE_ARRAYS(mask, nb, ix, iy, iz);
  const char* who = "lsm3ddeterminenarrowbandfrommask_";
  const Box B = LSMHIP_BOX(mg), nbb = LSMHIP_BOX(nbb);
  if (!check_box(nbb, B, who, "narrow_band") || !check_size(B, who) || !check_size(nbb, who)) return;
  n_lo[0] = *nlo;
  Level0 a{nullptr, mask, B.volume(), B.nx(), B.ny(), B, real_t(0), real_t(0), *use_mask_sign > 0 ? 1 : 2};
  Level0Out o{ix, iy, iz, *nlo, nullptr, 0, 0, nb, nbb};
  long long tot[3] = {0, 0, 0};
  if (!B.empty()) {
    if (!level0(who, a, o, tot)) return;
    if (tot[0] > (long long)*nhi - *nlo + 1) {
      set_error(kErrArg, "%s: %lld band points exceed the index arrays", who, tot[0]);
      return;
    }
    if (!level0_write(who, a, o)) return;
  }
  n_hi[0] = *nlo + (int)tot[0] - 1;
  grow_levels(who, nb, nbb, ix, iy, iz, *nlo, *nhi, n_lo, n_hi, *level);
  finish(who);
  take_device_flag(who);
}

namespace {
__global__ void __launch_bounds__(256) boundary_layer_kernel(unsigned char* nb, Box gb, Box fb, unsigned char mark,
                                                             int d) {
  // plane d/2 of the fill box (low side for even d), spanning the ghost box in the other axes
  const int ax = d / 2, hi = d % 2;
  const int a = (ax + 1) % 3, c = (ax + 2) % 3;
  const int lo_g[3] = {gb.ilo, gb.jlo, gb.klo}, hi_g[3] = {gb.ihi, gb.jhi, gb.khi};
  const int na = hi_g[a] - lo_g[a] + 1, nc = hi_g[c] - lo_g[c] + 1;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long long)na * nc) return;
  int q[3];
  q[ax] = hi ? (ax == 0 ? fb.ihi : ax == 1 ? fb.jhi : fb.khi) : (ax == 0 ? fb.ilo : ax == 1 ? fb.jlo : fb.klo);
  q[a] = lo_g[a] + (int)(t % na);
  q[c] = lo_g[c] + (int)(t / na);
  nb[(q[0] - gb.ilo) + gb.nx() * ((q[1] - gb.jlo) + gb.ny() * (q[2] - gb.klo))] = mark;
}
}  // namespace

// The 2*nd fill-box planes (lines in 2D), each spanning the ghost box of the other axes.
static void mark_boundary(const char* who, unsigned char* nb, const Box& gb, const Box& fb, unsigned char mark,
                          int nd) {
  if (gb.empty() || fb.empty()) return;
  // every plane index must lie in the ghost box (the reference writes them unchecked)
  if (fb.ilo < gb.ilo || fb.ihi > gb.ihi || fb.jlo < gb.jlo || fb.jhi > gb.jhi || fb.klo < gb.klo || fb.khi > gb.khi) {
    check_box(gb, fb, who, "narrow_band");
    return;
  }
  if (!check_size(gb, who)) return;
  for (int d = 0; d < 2 * nd; d++) {
    const int ax = d / 2;
    const long long n = (long long)(ax == 0 ? gb.ny() * gb.nz() : ax == 1 ? gb.nx() * gb.nz() : gb.nx() * gb.ny());
    hipLaunchKernelGGL(boundary_layer_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream(), nb, gb, fb,
                       mark, d);
  }
  finish(who);
}

// :303-377 -- six fill-box planes, each spanning the ghost box of the other axes.
LSMHIP_EXPORT void lsm3dmarknarrowbandboundarylayer_(unsigned char* nb, LSMHIP_BOXP(gb), LSMHIP_BOXP(fb),
                                                     const unsigned char* mark) {
  LSMHIP_REQUIRE_ARRAYS(nb);
  mark_boundary("lsm3dmarknarrowbandboundarylayer_", nb, LSMHIP_BOX(gb), LSMHIP_BOX(fb), *mark, 3);
}

namespace {
__global__ void __launch_bounds__(256) outer_sign_kernel(const real_t* phi, Box g, const int* ix, const int* iy,
                                                         const int* iz, int nlo, int nhi, const int* outer,
                                                         int nlo_outer, int lo, int hi, int* bits, int which,
                                                         int* flag) {
  const int m = lo + (int)(blockIdx.x * 256 + threadIdx.x);
  if (m > hi) return;
  const int l = outer[m - nlo_outer];
  if (l < nlo || l > nhi) {
    *flag = 1;
    return;
  }
  const int i = ix[l - nlo], j = iy[l - nlo], k = iz ? iz[l - nlo] : 0;
  if (!g.has(i, j, k)) {
    *flag = 1;
    return;
  }
  const real_t v = phi[(i - g.ilo) + g.nx() * ((j - g.jlo) + g.ny() * (k - g.klo))];
  atomicOr(&bits[which], v > real_t(0) ? 1 : 2);
}
}  // namespace

// :568-653 -- change_sign = 1 when the outer plus list or the outer minus list
// holds points of both signs (the reference's early exits do not change this).
static void check_outer(const char* who, int* change_sign, const real_t* phi, const Box& g, const int* ix,
                        const int* iy, const int* iz, const int* nlo, const int* nhi, const int* outer,
                        const int* nlo_outer, const int* nhi_outer, const int* nlo_plus, const int* nhi_plus,
                        const int* nlo_minus, const int* nhi_minus) {
  *change_sign = 0;
  int* bits = (int*)workspace(2 * sizeof(int));
  int* flag = device_flag();
  if (!bits || !flag) return;
  if (!check_hip(hipMemsetAsync(bits, 0, 2 * sizeof(int), stream()), who)) return;
  const int lo[2] = {*nlo_plus, *nlo_minus}, hi[2] = {*nhi_plus, *nhi_minus};
  for (int w = 0; w < 2; w++) {
    if (hi[w] < lo[w]) continue;
    if (lo[w] < *nlo_outer || hi[w] > *nhi_outer) {
      set_error(kErrArg, "%s: outer range [%d,%d] outside index_outer", who, lo[w], hi[w]);
      return;
    }
    const long long n = (long long)hi[w] - lo[w] + 1;
    hipLaunchKernelGGL(outer_sign_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream(), phi, g, ix, iy,
                       iz, *nlo, *nhi, outer, *nlo_outer, lo[w], hi[w], bits, w, flag);
  }
  int h[2] = {0, 0};
  if (!check_hip(hipMemcpyAsync(h, bits, sizeof(h), hipMemcpyDeviceToHost, stream()), who)) return;
  if (!check_hip(hipStreamSynchronize(stream()), who)) return;
  take_device_flag(who);
  *change_sign = (h[0] == 3 || h[1] == 3) ? 1 : 0;
}

LSMHIP_EXPORT void lsm3dcheckouternarrowbandlayer_(int* change_sign, const real_t* phi, LSMHIP_BOXP(g), const int* ix,
                                                   const int* iy, const int* iz, const int* nlo, const int* nhi,
                                                   const int* outer, const int* nlo_outer, const int* nhi_outer,
                                                   const int* nlo_plus, const int* nhi_plus, const int* nlo_minus,
                                                   const int* nhi_minus) {
  LSMHIP_REQUIRE_ARRAYS(phi, ix, iy, iz, outer);
  check_outer("lsm3dcheckouternarrowbandlayer_", change_sign, phi, LSMHIP_BOX(g), ix, iy, iz, nlo, nhi, outer,
              nlo_outer, nhi_outer, nlo_plus, nhi_plus, nlo_minus, nhi_minus);
}

// ---------------------------------------------------------------------------
// src/toolbox/lsm_localization2d.h entry points (2D boxes of four ints).
// ---------------------------------------------------------------------------
#define LSMHIP_BOXP2(n) const int *ilo_##n, const int *ihi_##n, const int *jlo_##n, const int *jhi_##n
#define LSMHIP_BOX2(n) ::lsmhip::Box{*ilo_##n, *ihi_##n, *jlo_##n, *jhi_##n, 0, 0}

// lsm2dDetermineNarrowBand, lsm_localization2d.f:139-243.  Unlike the 3D
// routine, the counters are stored whatever the band holds (n_hi(0) = nlo - 1
// for an empty band), and levels 1..level grow only when level 0 holds at
// least two points (:231); otherwise n_lo/n_hi(1..level) are left as they were.
LSMHIP_EXPORT void lsm2ddeterminenarrowband_(const real_t* phi, LSMHIP_BOXP2(g), unsigned char* nb, LSMHIP_BOXP2(nbb),
                                             int* ix, int* iy, const int* nlo, const int* nhi, int* n_lo, int* n_hi,
                                             int* outer, const int* nlo_outer, const int* nhi_outer,
                                             int* nlo_outer_plus, int* nhi_outer_plus, int* nlo_outer_minus,
                                             int* nhi_outer_minus, const real_t* width, const real_t* width_inner,
                                             const int* level) {
  LSMHIP_REQUIRE_ARRAYS(phi, nb, ix, iy, outer);
  const char* who = "lsm2ddeterminenarrowband_";
  const Box B = LSMHIP_BOX2(g), nbb = LSMHIP_BOX2(nbb);
  if (!check_box(nbb, B, who, "narrow_band") || !check_size(B, who) || !check_size(nbb, who)) return;
  n_lo[0] = *nlo;
  *nlo_outer_minus = *nlo_outer;
  *nhi_outer_plus = *nhi_outer;
  Level0 a{phi, nullptr, B.volume(), B.nx(), B.ny(), B, *width, *width_inner, 0};
  Level0Out o{ix, iy, nullptr, *nlo, outer, *nlo_outer, *nhi_outer, nb, nbb};
  long long tot[3] = {0, 0, 0};
  if (!B.empty()) {
    if (!level0(who, a, o, tot)) return;
    if (tot[0] > (long long)*nhi - *nlo + 1 || tot[1] + tot[2] > (long long)*nhi_outer - *nlo_outer + 1) {
      set_error(kErrArg, "%s: %lld band points (%lld outer) exceed the index arrays", who, tot[0], tot[1] + tot[2]);
      return;
    }
    if (!level0_write(who, a, o)) return;
  }
  n_hi[0] = *nlo + (int)tot[0] - 1;
  *nhi_outer_minus = *nlo_outer + (int)tot[1] - 1;
  *nlo_outer_plus = *nhi_outer - (int)tot[2] + 1;
  if (*level > 0 && n_hi[0] > n_lo[0]) grow_levels(who, nb, nbb, ix, iy, nullptr, *nlo, *nhi, n_lo, n_hi, *level);
  finish(who);
  take_device_flag(who);
}

// lsm2dDetermineNarrowBandFromMask, :539-615: level 0 is {mask >= 0} when
// use_mask_sign > 0, else {mask < 0} (strict; the 3D routine uses <= 0).
LSMHIP_EXPORT void lsm2ddeterminenarrowbandfrommask_(const real_t* mask, LSMHIP_BOXP2(mg), unsigned char* nb,
                                                     LSMHIP_BOXP2(nbb), int* ix, int* iy, const int* nlo,
                                                     const int* nhi, int* n_lo, int* n_hi, const int* level,
                                                     const int* use_mask_sign) {
  LSMHIP_REQUIRE_ARRAYS(mask, nb, ix, iy);
  const char* who = "lsm2ddeterminenarrowbandfrommask_";
  const Box B = LSMHIP_BOX2(mg), nbb = LSMHIP_BOX2(nbb);
  if (!check_box(nbb, B, who, "narrow_band") || !check_size(B, who) || !check_size(nbb, who)) return;
  n_lo[0] = *nlo;
  Level0 a{nullptr, mask, B.volume(), B.nx(), B.ny(), B, real_t(0), real_t(0), *use_mask_sign > 0 ? 1 : 3};
  Level0Out o{ix, iy, nullptr, *nlo, nullptr, 0, 0, nb, nbb};
  long long tot[3] = {0, 0, 0};
  if (!B.empty()) {
    if (!level0(who, a, o, tot)) return;
    if (tot[0] > (long long)*nhi - *nlo + 1) {
      set_error(kErrArg, "%s: %lld band points exceed the index arrays", who, tot[0]);
      return;
    }
    if (!level0_write(who, a, o)) return;
  }
  n_hi[0] = *nlo + (int)tot[0] - 1;
  grow_levels(who, nb, nbb, ix, iy, nullptr, *nlo, *nhi, n_lo, n_hi, *level);
  finish(who);
  take_device_flag(who);
}

// :486-533 -- the four fill-box lines, each spanning the ghost box.
LSMHIP_EXPORT void lsm2dmarknarrowbandboundarylayer_(unsigned char* nb, LSMHIP_BOXP2(gb), LSMHIP_BOXP2(fb),
                                                     const unsigned char* mark) {
  LSMHIP_REQUIRE_ARRAYS(nb);
  mark_boundary("lsm2dmarknarrowbandboundarylayer_", nb, LSMHIP_BOX2(gb), LSMHIP_BOX2(fb), *mark, 2);
}

// :693-772
LSMHIP_EXPORT void lsm2dcheckouternarrowbandlayer_(int* change_sign, const real_t* phi, LSMHIP_BOXP2(g), const int* ix,
                                                   const int* iy, const int* nlo, const int* nhi, const int* outer,
                                                   const int* nlo_outer, const int* nhi_outer, const int* nlo_plus,
                                                   const int* nhi_plus, const int* nlo_minus, const int* nhi_minus) {
  LSMHIP_REQUIRE_ARRAYS(phi, ix, iy, outer);
  check_outer("lsm2dcheckouternarrowbandlayer_", change_sign, phi, LSMHIP_BOX2(g), ix, iy, nullptr, nlo, nhi, outer,
              nlo_outer, nhi_outer, nlo_plus, nhi_plus, nlo_minus, nhi_minus);
}
