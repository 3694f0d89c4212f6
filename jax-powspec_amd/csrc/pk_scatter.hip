// Particle -> mesh mass assignment: reference stencil (src/mas.py:88-153 and :5-87),
// textbook CIC and TSC.
//
// Sorted owner-computes deposit on row-bundle tiles (1 x ty x n cells, full z rows):
//   1. count:  each block takes a chunk of particles (one block per CU) and histograms their
//              keys (tile, y-class) in LDS, or in its row of the count matrix in HBM when
//              the keys do not fit; the block's row of counts goes to a [blocks][keys]
//              matrix.
//   2. scan:   a column scan of the matrix gives every (block, key) its offset inside the
//              key; an exclusive scan of the key totals gives the key starts.
//   3. place, in two passes.  Keys are x-plane major, so the records of one x plane fill
//              one contiguous range of the sorted order.
//      a. coarse: the blocks re-read their chunk and write each record into its x plane's
//              range (block offset = sum of the block's key offsets over the plane, plus an
//              LDS rank).  A block appends to at most n + 1 runs, so the partly written
//              128-byte lines of all blocks fit the L2 and leave it whole.
//      b. fine:  blocks take consecutive segments of the plane-grouped records, histogram
//              their keys over the segment's window of planes in LDS, reserve one run per
//              key with a global atomic and write each key's records as one run.
//              Owner-computes deposits sum in fixed point, so the order inside a key
//              (atomic order) does not change any result.
//   4. deposit: one block per tile stages its ty full rows in LDS (zeros or mesh_in), walks
//              its own buckets plus the neighbour buckets that spill into it (the whole
//              x - 1 / x + 1 tiles, the boundary-row classes of the y neighbours), adds every
//              stencil corner that falls inside with LDS atomics (64-bit fixed point) and
//              writes each row once with coalesced stores: no global atomics, no zero fill.
//   5. atomic: the (normally empty) irregular bucket is added with global atomics.
// A direct mode (tile[0] == 0) adds every particle with global atomics instead.
#include <mutex>
#include <utility>
#include <vector>

#include "pk_particles.h"

namespace pk {

// ------------------------------------------------------------------------------------
// small utility kernels (fills instead of hipMemset so a captured graph has no memset node)
// ------------------------------------------------------------------------------------
__global__ void k_fill_f64(double* __restrict__ p, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_copy_or_zero(const float* src, float* dst, int64_t n) {  // src may alias dst
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src ? src[i] : 0.0f;
}

constexpr int CNT_NT = 1024;
constexpr int CNT_UNROLL = 8;

// Batched wrappers: blockIdx.y selects the job; blocks beyond the job's own extent return
// at once (the whole block, so no barrier is left waiting).
#define PK_JOB_OR_RETURN(jobs, limit_expr)                              \
  const Job& j = (jobs)[blockIdx.y];                                    \
  if (!j.active || (int)blockIdx.x >= (limit_expr)) return;

// ------------------------------------------------------------------------------------
// 1. count: LDS histogram of one chunk -> matrix row; FP64 weight sum per block
// ------------------------------------------------------------------------------------
struct CountShared {
  double red[CNT_NT];
  unsigned wmax_bits;
};

// AOS / W: the job's input layout and weights, fixed per instantiation (no branch between the
// loads of a thread's particles)
template <int AOS, int W, int ST = -1>
__device__ __forceinline__ void bucket_count_body(const Job& j, int bx, CountShared& sh) {
  // chunks of <= 65535 particles: key k -> 16-bit half (k & 1) of word k >> 1 (half the
  // LDS); larger chunks: one 32-bit counter per key
  extern __shared__ unsigned hist[];
  double* red = sh.red;
  unsigned& wmax_bits = sh.wmax_bits;
  const ScatterArgs a = j.a;
  const TileGeom g = j.g;
  const int tid = threadIdx.x;
  const bool packed = g.chunk <= 65535;
  unsigned* row = j.mat + (size_t)bx * g.nkeys;
  if (g.gtab) {   // per-key counters in the block's own row of the count matrix (HBM)
    for (int k = tid; k < g.nkeys; k += CNT_NT) row[k] = 0u;
    __threadfence();   // the zeros reach L2, where the atomics below add, before any adds
  } else {
    const int words = packed ? (g.nkeys + 1) >> 1 : g.nkeys;
    for (int i = tid; i < words; i += CNT_NT) hist[i] = 0u;
  }
  if (tid == 0) wmax_bits = 0u;
  // block 0 clears the estimate's statistics (written from the scans on; no separate fill)
  if (bx == 0 && tid < PK_STATS_LEN) j.stats[tid] = 0.0;
  __syncthreads();
  const int64_t lo = (int64_t)bx * g.chunk;
  const int64_t hi = min(lo + (int64_t)g.chunk, a.np);
  double ws = 0.0;
  float wm = 0.0f;
  // CNT_UNROLL particles per thread and iteration: their loads are all in flight before
  // the first key is computed
  for (int64_t i0 = lo + tid; i0 < hi; i0 += (int64_t)CNT_NT * CNT_UNROLL) {
    float px[CNT_UNROLL], py[CNT_UNROLL], pz[CNT_UNROLL], w[CNT_UNROLL];
    // indexes clamped to the chunk: the loads need no branch (the extra ones repeat the
    // last particle and are not counted)
#pragma unroll
    for (int u = 0; u < CNT_UNROLL; ++u)
      load_particle<AOS, W, true>(a, min(i0 + (int64_t)u * CNT_NT, hi - 1), px[u], py[u], pz[u], w[u]);
    __builtin_amdgcn_sched_barrier(0);   // every load issued before the first key
#pragma unroll
    for (int u = 0; u < CNT_UNROLL; ++u) {
      if (i0 + (int64_t)u * CNT_NT >= hi) continue;
      const int key = row_tile_key<ST>(a, g, px[u], py[u], pz[u]);
      if (key >= 0) {
        if (g.gtab) atomicAdd(&row[key], 1u);
        else if (packed) atomicAdd(&hist[key >> 1], 1u << (16 * (key & 1)));
        else atomicAdd(&hist[key], 1u);
        ws += (double)w[u];
        wm = fmaxf(wm, fabsf(w[u]));
      }
    }
  }
  atomicMax(&wmax_bits, __float_as_uint(wm));   // non-negative floats order like their bits
  __syncthreads();
  if (tid == 0) j.wmax[bx] = (double)__uint_as_float(wmax_bits);
  if (!g.gtab)
    for (int k = tid; k < g.nkeys; k += CNT_NT) row[k] = packed ? (hist[k >> 1] >> (16 * (k & 1))) & 0xffffu : hist[k];
  const double t = block_sum_f64(ws, red, CNT_NT);
  if (tid == 0) j.wsum[bx] = t;
}
template <int AOS, int W, int ST>
__global__ void __launch_bounds__(CNT_NT) k_bucket_count(Job j) {
  __shared__ CountShared sh;
  bucket_count_body<AOS, W, ST>(j, blockIdx.x, sh);
}
__global__ void __launch_bounds__(CNT_NT) k_bucket_count_jobs(const Job* __restrict__ jobs) {
  __shared__ CountShared sh;
  PK_JOB_OR_RETURN(jobs, j.g.nblk);
  // block-uniform choice of the instantiation for this job's layout and weights
  const bool w = j.a.w != nullptr;
  if (j.a.aos) {
    if (w) bucket_count_body<1, 1>(j, blockIdx.x, sh);
    else bucket_count_body<1, 0>(j, blockIdx.x, sh);
  } else {
    if (w) bucket_count_body<0, 1>(j, blockIdx.x, sh);
    else bucket_count_body<0, 0>(j, blockIdx.x, sh);
  }
}

// ------------------------------------------------------------------------------------
// 2. column scan (offset of each block inside each key) + exclusive scan of key totals
// ------------------------------------------------------------------------------------
constexpr int COL_NT = 256;
constexpr int COL_KEYS = 32;                  // keys per block: one 128-byte segment per row
constexpr int COL_SUBS = COL_NT / COL_KEYS;   // threads per key, each over a run of rows
constexpr int COL_ROWS = 32;                  // rows per thread and pass (loads in flight)

// Thread (sub, key) walks rows [sub * COL_ROWS, +COL_ROWS) of each pass of COL_SUBS *
// COL_ROWS rows: the run sums of a key's threads meet in LDS, so one pass costs one round
// trip to memory instead of a walk down the whole column.
__device__ __forceinline__ void col_scan_body(const Job& j, int bx) {
  __shared__ unsigned part[COL_SUBS][COL_KEYS];
  const int nkeys = j.g.nkeys, nblk = j.g.nblk;
  unsigned* mat = j.mat;
  const int kk = threadIdx.x % COL_KEYS, sub = threadIdx.x / COL_KEYS;
  const int k = bx * COL_KEYS + kk;
  const bool on = k < nkeys;
  unsigned carry = 0;
  for (int r0 = 0; r0 < nblk; r0 += COL_SUBS * COL_ROWS) {   // block-uniform
    const int rb = r0 + sub * COL_ROWS;
    unsigned c[COL_ROWS];
    unsigned sum = 0;
#pragma unroll
    for (int u = 0; u < COL_ROWS; ++u) {
      c[u] = (on && rb + u < nblk) ? mat[(size_t)(rb + u) * nkeys + k] : 0u;
      sum += c[u];
    }
    part[sub][kk] = sum;
    __syncthreads();
    unsigned run = carry, total = 0;
#pragma unroll
    for (int t = 0; t < COL_SUBS; ++t) {
      const unsigned v = part[t][kk];
      if (t < sub) run += v;
      total += v;
    }
    __syncthreads();   // part is reused by the next pass
#pragma unroll
    for (int u = 0; u < COL_ROWS; ++u) {
      if (on && rb + u < nblk) {
        mat[(size_t)(rb + u) * nkeys + k] = run;
        run += c[u];
      }
    }
    carry += total;
  }
  if (on && sub == 0) {
    j.totals[k] = (int)carry;
    j.koff[k] = 0u;   // the fine placing pass counts each key's records from zero
  }
}
__global__ void __launch_bounds__(COL_NT) k_col_scan(Job j) { col_scan_body(j, blockIdx.x); }
__global__ void __launch_bounds__(COL_NT) k_col_scan_jobs(const Job* __restrict__ jobs) {
  PK_JOB_OR_RETURN(jobs, (j.g.nkeys + COL_KEYS - 1) / COL_KEYS);
  col_scan_body(j, blockIdx.x);
}

constexpr int SCAN_NT = 1024;
constexpr int SCAN_PER = 4;
constexpr int SCAN_BLOCK = SCAN_NT * SCAN_PER;

// wave scans with shuffles, then the wave totals (two barriers)
__device__ int block_exclusive_scan(int v, int* sh, int& total) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  int incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  if (lane == 63) sh[w] = incl;
  __syncthreads();
  int before = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < SCAN_NT / 64; ++i) {
    const int x = sh[i];
    if (i < w) before += x;
    tot += x;
  }
  total = tot;
  __syncthreads();
  return before + incl - v;
}

// key totals -> key starts: local scans of SCAN_BLOCK keys, a scan of the block sums, add
__device__ __forceinline__ void scan_local_body(const Job& j, int bx) {
  __shared__ int sh[SCAN_NT];
  const int n = j.g.nkeys;
  const int* in = j.totals;
  int* out = j.key_start;
  const int base = bx * SCAN_BLOCK + threadIdx.x * SCAN_PER;
  int v[SCAN_PER];
  int s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  int total;
  int run = block_exclusive_scan(s, sh, total);
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) j.bsum[bx] = total;
}
__global__ void __launch_bounds__(SCAN_NT) k_scan_local(Job j) { scan_local_body(j, blockIdx.x); }
__global__ void __launch_bounds__(SCAN_NT) k_scan_local_jobs(const Job* __restrict__ jobs) {
  PK_JOB_OR_RETURN(jobs, (j.g.nkeys + SCAN_BLOCK - 1) / SCAN_BLOCK);
  scan_local_body(j, blockIdx.x);
}

// key_start[nkeys] = the records' total; sums the weight partials into stats[3] =
// sum(w) / n^3 and sets the fixed-point scale aux[0] = S, aux[1] = 1 / S from max |w|.
// Every thread of the block must call it.
__device__ __forceinline__ void scan_finish_body(const Job& j, int total) {
  __shared__ double shd[SCAN_NT];
  __shared__ unsigned long long wmax_bits;
  if (threadIdx.x == 0) j.key_start[j.g.nkeys] = total;
  double ws = 0.0, wm = 0.0;
  if (threadIdx.x == 0) wmax_bits = 0ull;
  for (int i = threadIdx.x; i < j.g.nblk; i += SCAN_NT) {
    ws += j.wsum[i];
    wm = fmax(wm, j.wmax[i]);
  }
  const double t = block_sum_f64_waves<SCAN_NT>(ws, shd);   // ends with a barrier
  atomicMax(&wmax_bits, (unsigned long long)__double_as_longlong(wm));
  __syncthreads();
  if (threadIdx.x == 0) {
    const double n = (double)j.a.n;
    j.stats[3] = t / (n * n * n);
    // fixed-point scale S = 2^30 / 2^e with max|w| < 2^e: every deposit |rint(w W S)| < 2^30
    // for corner weights W <= 1, so it converts through int32
    int e = 0;
    const double wmax = __longlong_as_double((long long)wmax_bits);
    if (wmax > 0.0 && wmax < 1e300) (void)frexp(wmax, &e);
    j.aux[0] = ldexp(1.0, 30 - e);
    j.aux[1] = ldexp(1.0, e - 30);
  }
}
// all keys in one local block: the local scan, then the finish (no add pass follows)
__global__ void __launch_bounds__(SCAN_NT) k_scan_single(Job j) {
  scan_local_body(j, 0);
  __syncthreads();
  scan_finish_body(j, j.bsum[0]);
}

// Block bx adds the sum of the block sums before it to its keys' local starts (each block
// sums that short prefix itself, so no separate scan of the block sums runs); block 0 also
// finishes (the records' total and the weight statistics).
__device__ __forceinline__ void scan_add_body(const Job& j, int bx) {
  __shared__ int sh[SCAN_NT / 64];
  const int n = j.g.nkeys;
  const int nb2 = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
  const int* bsum = j.bsum;
  const int upto = bx == 0 ? nb2 : bx;   // block 0: every block sum (the total); else those before bx
  int s = 0;
  for (int i = threadIdx.x; i < upto; i += SCAN_NT) s += bsum[i];
  int sum;
  (void)block_exclusive_scan(s, sh, sum);
  int* out = j.key_start;
  const int base = bx * SCAN_BLOCK + threadIdx.x * SCAN_PER;
  const int add = bx == 0 ? 0 : sum;
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k)
    if (base + k < n) out[base + k] += add;
  if (bx == 0) scan_finish_body(j, sum);   // block-uniform
}
__global__ void __launch_bounds__(SCAN_NT) k_scan_add(Job j) { scan_add_body(j, blockIdx.x); }
__global__ void __launch_bounds__(SCAN_NT) k_scan_add_jobs(const Job* __restrict__ jobs) {
  PK_JOB_OR_RETURN(jobs, (j.g.nkeys + SCAN_BLOCK - 1) / SCAN_BLOCK);
  scan_add_body(j, blockIdx.x);
}

// ------------------------------------------------------------------------------------
// 3. place (both passes): rounds of records are ranked per destination run in LDS, staged
// there in destination order and written by consecutive lanes to consecutive addresses, so
// a wave's stores touch a few runs instead of one line per lane
// ------------------------------------------------------------------------------------

// Exclusive scan of the counts a[0, len) in LDS by the whole block of NT threads (each takes
// a contiguous span): f(i, count, prefix) for every entry, in the thread that owns it.
// Returns the grand total.  Every thread must call it (it holds a barrier).
template <int NT, class F>
__device__ __forceinline__ unsigned block_scan_apply(const unsigned* a, int len, unsigned* wtot, F&& f) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int per = (len + NT - 1) / NT;
  const int i0 = min(tid * per, len), i1 = min(i0 + per, len);
  unsigned own = 0;
  for (int i = i0; i < i1; ++i) own += a[i];
  unsigned incl = own;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned v = __shfl_up(incl, off, 64);
    if (lane >= off) incl += v;
  }
  if (lane == 63) wtot[wv] = incl;
  __syncthreads();
  unsigned pre = 0, total = 0;
#pragma unroll
  for (int k = 0; k < NT / 64; ++k) {
    const unsigned t = wtot[k];
    if (k < wv) pre += t;
    total += t;
  }
  unsigned e = pre + incl - own;
  for (int i = i0; i < i1; ++i) {
    const unsigned c = a[i];
    f(i, c, e);
    e += c;
  }
  return total;
}

// ------------------------------------------------------------------------------------
// 3a. coarse place: x-plane start + block offset inside the plane + LDS rank
// ------------------------------------------------------------------------------------
constexpr int PLACE_PER = 7;                     // most particles per thread and round
constexpr size_t PLACE_LDS = 150 * 1024;         // tables + staging of the coarse pass
// particles per round: 20 staged bytes each (x, y, z, w, destination) next to three tables
// of n + 1 words, a multiple of the block size
__host__ __device__ constexpr int place_round(int n) {
  return (int)((PLACE_LDS - 12 * (size_t)(n + 1)) / 20) / CNT_NT * CNT_NT > CNT_NT * PLACE_PER
             ? CNT_NT * PLACE_PER
             : (int)((PLACE_LDS - 12 * (size_t)(n + 1)) / 20) / CNT_NT * CNT_NT;
}
size_t place_lds(int n) { return 12 * ((size_t)n + 1) + 20 * (size_t)place_round(n); }

template <int AOS, int W, int ST = -1>
__device__ __forceinline__ void bucket_place_body(const Job& j, int bx, unsigned* wtot) {
  extern __shared__ unsigned lds[];
  const ScatterArgs a = j.a;
  const TileGeom g = j.g;
  const int tid = threadIdx.x;
  const int G = a.n + 1;            // x planes, then the irregular key
  const int S = place_round(a.n);   // particles per round
  unsigned* dest = lds;             // next write position of each plane for this block
  unsigned* lcnt = lds + G;         // the round's count, then local start, of each plane
  unsigned* delta = lds + 2 * G;    // global position - local slot, per plane
  float* sx = reinterpret_cast<float*>(lds + 3 * G);
  float* sy = sx + S;
  float* sz = sy + S;
  float* sw = sz + S;
  unsigned* sd = reinterpret_cast<unsigned*>(sw + S);   // global position of each slot
  const unsigned kpp = (unsigned)(g.nyt * g.ncls);
  {
    // plane start + the records of this plane in earlier blocks (the sum of this block's
    // offsets over the plane's keys): one wave per plane
    const unsigned* row = j.mat + (size_t)bx * g.nkeys;
    const int wv = tid >> 6, ln = tid & 63;
    for (int i = wv; i < G; i += CNT_NT / 64) {
      const int k0 = i * (int)kpp, k1 = i < a.n ? k0 + (int)kpp : g.nkeys;
      unsigned sum = 0;
      for (int k = k0 + ln; k < k1; k += 64) sum += row[k];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
      if (ln == 0) dest[i] = (unsigned)j.key_start[k0] + sum;
    }
  }
  const int64_t lo = (int64_t)bx * g.chunk;
  const int64_t hi = min(lo + (int64_t)g.chunk, a.np);
  const bool has_w = a.w != nullptr;
  // software pipeline: the next round's particles load while this round is ranked, staged
  // and stored
  float px[PLACE_PER], py[PLACE_PER], pz[PLACE_PER], w[PLACE_PER];
  // indexes clamped to the chunk: no branch between the loads (the extra loads repeat the
  // chunk's last particle and are not placed)
  auto load_round = [&](int64_t r0, float* x, float* y, float* z, float* ww) {
#pragma unroll
    for (int u = 0; u < PLACE_PER; ++u)
      load_particle<AOS, W, true>(a, min(r0 + (int64_t)u * CNT_NT + tid, hi - 1), x[u], y[u], z[u], ww[u]);
  };
  if (lo >= hi) return;   // whole block: an empty chunk
  load_round(lo, px, py, pz, w);
  for (int64_t r0 = lo; r0 < hi; r0 += S) {   // block-uniform
    int pl[PLACE_PER];
#pragma unroll
    for (int u = 0; u < PLACE_PER; ++u) {
      const int64_t i = r0 + (int64_t)u * CNT_NT + tid;
      pl[u] = (u * CNT_NT < S && i < hi) ? row_plane<ST>(a, px[u], py[u], pz[u]) : -1;
    }
    float nx[PLACE_PER], ny[PLACE_PER], nz[PLACE_PER], nw[PLACE_PER];
    load_round(r0 + S, nx, ny, nz, nw);
    for (int q = tid; q < G; q += CNT_NT) lcnt[q] = 0u;
    __syncthreads();
    unsigned rk[PLACE_PER];
#pragma unroll
    for (int u = 0; u < PLACE_PER; ++u)
      if (pl[u] >= 0) rk[u] = atomicAdd(&lcnt[pl[u]], 1u);
    __syncthreads();
    const unsigned m = block_scan_apply<CNT_NT>(lcnt, G, wtot, [&](int p, unsigned c, unsigned e) {
      lcnt[p] = e;
      if (c) {   // this thread alone owns plane p here
        const unsigned b = dest[p];
        dest[p] = b + c;
        delta[p] = b - e;
      }
    });
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PLACE_PER; ++u)
      if (pl[u] >= 0) {
        const unsigned slot = lcnt[pl[u]] + rk[u];
        sx[slot] = px[u];
        sy[slot] = py[u];
        sz[slot] = pz[u];
        sw[slot] = w[u];
        sd[slot] = delta[pl[u]] + slot;
      }
    __syncthreads();
    for (int q = tid; q < (int)m; q += CNT_NT) store_record(j.staged, sd[q], has_w, sx[q], sy[q], sz[q], sw[q]);
#pragma unroll
    for (int u = 0; u < PLACE_PER; ++u) {
      px[u] = nx[u];
      py[u] = ny[u];
      pz[u] = nz[u];
      w[u] = nw[u];
    }
    __syncthreads();   // the staging is read before the next round overwrites it
  }
}
template <int AOS, int W, int ST>
__global__ void __launch_bounds__(CNT_NT) k_bucket_place(Job j) {
  __shared__ unsigned wtot[CNT_NT / 64];
  bucket_place_body<AOS, W, ST>(j, blockIdx.x, wtot);
}
__global__ void __launch_bounds__(CNT_NT) k_bucket_place_jobs(const Job* __restrict__ jobs) {
  __shared__ unsigned wtot[CNT_NT / 64];
  PK_JOB_OR_RETURN(jobs, j.g.nblk);
  const bool w = j.a.w != nullptr;
  if (j.a.aos) {
    if (w) bucket_place_body<1, 1>(j, blockIdx.x, wtot);
    else bucket_place_body<1, 0>(j, blockIdx.x, wtot);
  } else {
    if (w) bucket_place_body<0, 1>(j, blockIdx.x, wtot);
    else bucket_place_body<0, 0>(j, blockIdx.x, wtot);
  }
}

// the plain launches' instantiation for a job's layout, weights and stencil
template <template <int, int, int> class Pick, int AOS, int W>
auto pick_st(const ScatterArgs& a) -> decltype(Pick<0, 0, 0>::k) {
  if (a.stencil == PK_STENCIL_CIC_REF) return Pick<AOS, W, PK_STENCIL_CIC_REF>::k;
  if (a.stencil == PK_STENCIL_CIC) return Pick<AOS, W, PK_STENCIL_CIC>::k;
  return Pick<AOS, W, PK_STENCIL_TSC>::k;
}
template <template <int, int, int> class Pick>
auto pick_aw(const ScatterArgs& a) -> decltype(Pick<0, 0, 0>::k) {
  const bool w = a.w != nullptr;
  if (a.aos) return w ? pick_st<Pick, 1, 1>(a) : pick_st<Pick, 1, 0>(a);
  return w ? pick_st<Pick, 0, 1>(a) : pick_st<Pick, 0, 0>(a);
}
template <int AOS, int W, int ST> struct CountK { static constexpr void (*k)(Job) = k_bucket_count<AOS, W, ST>; };
template <int AOS, int W, int ST> struct PlaceK { static constexpr void (*k)(Job) = k_bucket_place<AOS, W, ST>; };

// ------------------------------------------------------------------------------------
// 3b. fine place: a segment of the plane-grouped records -> one run per (segment, key)
// ------------------------------------------------------------------------------------
constexpr int FINE_NT = 256;                      // small blocks, several per CU: their
constexpr int FINE_PER = 6;                       // load, rank and store phases overlap
constexpr int FINE_SEG = FINE_NT * FINE_PER;      // records per block
constexpr int FINE_WIN = 1024;                    // keys of the LDS window
constexpr size_t FINE_LDS = (size_t)FINE_SEG * 20;   // staging: x, y, z, w, destination

int fine_place_blocks(int64_t np) { return (int)((np > 0 ? np + FINE_SEG - 1 : FINE_SEG) / FINE_SEG); }

template <int ST = -1>
__device__ __forceinline__ void fine_place_body(const Job& j, int bx) {
  extern __shared__ unsigned lds[];
  __shared__ unsigned cnt[FINE_WIN];     // per key of the window: count, then local start
  __shared__ unsigned delta[FINE_WIN];   // global position - local slot
  __shared__ unsigned wtot[FINE_NT / 64];
  __shared__ int win[2];
  const ScatterArgs a = j.a;
  const TileGeom g = j.g;
  const int tid = threadIdx.x;
  const int64_t total = j.key_start[g.nkeys];   // records kept
  const int64_t s0 = (int64_t)bx * FINE_SEG;
  if (s0 >= total) return;   // whole block
  const int64_t s1 = min(s0 + (int64_t)FINE_SEG, total);
  const bool has_w = a.w != nullptr;
  const unsigned kpp = (unsigned)(g.nyt * g.ncls);
  float* sx = reinterpret_cast<float*>(lds);
  float* sy = sx + FINE_SEG;
  float* sz = sy + FINE_SEG;
  float* sw = sz + FINE_SEG;
  unsigned* sd = reinterpret_cast<unsigned*>(sw + FINE_SEG);
  // the records are grouped by plane, so the first and last records bound the segment's planes
  if (tid < 2) {
    const float4 r = load_record(j.staged, (size_t)(tid == 0 ? s0 : s1 - 1), has_w);
    int plane = 0;
    (void)row_tile_key<ST>(a, g, r.x, r.y, r.z, &plane);
    win[tid] = plane;
  }
  float4 r[FINE_PER];
  int key[FINE_PER];
#pragma unroll
  for (int u = 0; u < FINE_PER; ++u)   // clamped: no branch between the loads
    r[u] = load_record_stream(j.staged, (size_t)min(s0 + tid + (int64_t)u * FINE_NT, s1 - 1), has_w);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int u = 0; u < FINE_PER; ++u) {
    const int64_t i = s0 + tid + (int64_t)u * FINE_NT;
    r[u].w = record_weight(r[u], has_w);
    key[u] = i < s1 ? row_tile_key<ST>(a, g, r[u].x, r[u].y, r[u].z) : -1;   // the key the coarse pass saw
  }
  __syncthreads();
  const int kbase = win[0] * (int)kpp;
  const int wlen = min((win[1] + 1) * (int)kpp, g.nkeys) - kbase;
  if (wlen > FINE_WIN) {
    // a sparse segment spanning too many planes (small catalogues): one global atomic each
#pragma unroll
    for (int u = 0; u < FINE_PER; ++u)
      if (key[u] >= 0) {
        const unsigned at = (unsigned)j.key_start[key[u]] + atomicAdd(&j.koff[key[u]], 1u);
        store_record(j.sorted, at, has_w, r[u].x, r[u].y, r[u].z, r[u].w);
      }
    return;
  }
  for (int q = tid; q < wlen; q += FINE_NT) cnt[q] = 0u;
  __syncthreads();
  unsigned rank[FINE_PER];
#pragma unroll
  for (int u = 0; u < FINE_PER; ++u)
    if (key[u] >= 0) rank[u] = atomicAdd(&cnt[key[u] - kbase], 1u);
  __syncthreads();
  const unsigned m = block_scan_apply<FINE_NT>(cnt, wlen, wtot, [&](int q, unsigned c, unsigned e) {
    cnt[q] = e;
    if (c) delta[q] = (unsigned)j.key_start[kbase + q] + atomicAdd(&j.koff[kbase + q], c) - e;
  });
  __syncthreads();
#pragma unroll
  for (int u = 0; u < FINE_PER; ++u)
    if (key[u] >= 0) {
      const int q = key[u] - kbase;
      const unsigned slot = cnt[q] + rank[u];
      sx[slot] = r[u].x;
      sy[slot] = r[u].y;
      sz[slot] = r[u].z;
      sw[slot] = r[u].w;
      sd[slot] = delta[q] + slot;
    }
  __syncthreads();
  for (int q = tid; q < (int)m; q += FINE_NT) store_record(j.sorted, sd[q], has_w, sx[q], sy[q], sz[q], sw[q]);
}
template <int ST>
__global__ void __launch_bounds__(FINE_NT) k_fine_place(Job j) { fine_place_body<ST>(j, blockIdx.x); }
__global__ void __launch_bounds__(FINE_NT) k_fine_place_jobs(const Job* __restrict__ jobs) {
  const Job& j = jobs[blockIdx.y];
  if (!j.active) return;
  fine_place_body(j, blockIdx.x);
}

// ------------------------------------------------------------------------------------
// 4. deposit (mesh mode): one block per row-bundle tile
// ------------------------------------------------------------------------------------
constexpr int DEP_NT = 512;

template <bool FIXED>
__global__ void __launch_bounds__(DEP_NT) k_rows_deposit(ScatterArgs a, TileGeom g, const int* __restrict__ key_start,
                                                         const void* __restrict__ sorted, const double* __restrict__ aux,
                                                         const float* mesh_in, float* mesh_out,
                                                         double* __restrict__ tile_mass) {
  extern __shared__ unsigned long long acc_q[];   // [ty][n]: 64-bit fixed point, or float when !FIXED
  float* acc_f = reinterpret_cast<float*>(acc_q);
  __shared__ int s_key[MAX_TILE_BUCKETS], s_base[MAX_TILE_BUCKETS], s_pre[MAX_TILE_BUCKETS + 1];
  __shared__ int s_nbk;
  __shared__ double red[DEP_NT];
  const int tid = threadIdx.x;
  const int n = a.n;
  const int t = blockIdx.x;
  const int xt = t / g.nyt, yt = t % g.nyt;
  const int y0 = yt * g.ty;
  const int cells = g.ty * n;
  const size_t gbase = ((size_t)xt * n + y0) * n;   // rows y0 .. y0+ty-1 are contiguous
  for (int i = tid; i < cells; i += DEP_NT) {
    if (FIXED) acc_q[i] = 0ull;
    else acc_f[i] = 0.0f;
  }
  gather_buckets(g, n, a.stencil == PK_STENCIL_TSC, false, xt, yt, key_start, s_key, s_base, s_pre, &s_nbk);
  const float scale = (float)aux[0];
  deposit_buckets(a, g, sorted, s_base, s_pre, s_nbk, DEP_NT, xt, y0, [&](int ly, int iz, float wt) {
    if (FIXED) atomicAdd(&acc_q[ly * n + iz], to_fixed(wt, scale));
    else atomicAdd(&acc_f[ly * n + iz], wt);
  });
  __syncthreads();
  // mesh_in may alias mesh_out: each cell is read once and then written by this thread
  const double inv = aux[1];
  double msum = 0.0;
  for (int i = tid; i < cells; i += DEP_NT) {
    const float v = FIXED ? (float)((double)(long long)acc_q[i] * inv) : acc_f[i];
    msum += FIXED ? (double)(long long)acc_q[i] * inv : (double)v;
    mesh_out[gbase + i] = mesh_in ? mesh_in[gbase + i] + v : v;
  }
  const double m = block_sum_f64(msum, red, DEP_NT);
  if (tid == 0) tile_mass[t] = m;
}

// ------------------------------------------------------------------------------------
// 5. global-atomic deposit: all particles (direct mode) or the irregular bucket
// ------------------------------------------------------------------------------------
constexpr int ATM_NT = 256;

__global__ void __launch_bounds__(ATM_NT) k_atomic_deposit(ScatterArgs a, const void* __restrict__ sorted,
                                                           const int* __restrict__ range, float* __restrict__ mesh,
                                                           double* __restrict__ stats) {
  __shared__ double red[ATM_NT];
  int64_t lo = 0, hi = a.np;
  if (range) {
    lo = range[0];
    hi = range[1];
  }
  const int n = a.n;
  double msum = 0.0;
  for (int64_t i = lo + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < hi; i += (int64_t)gridDim.x * blockDim.x) {
    float px, py, pz, w;
    if (sorted) {
      const float4 r = load_record(sorted, (size_t)i, a.w != nullptr);
      px = r.x; py = r.y; pz = r.z; w = r.w;
    } else {
      load_particle(a, i, px, py, pz, w);
    }
    if (!pos_ok(px, py, pz)) continue;
    for_each_corner(a, px, py, pz, w, [&](int ix, int iy, int iz, float wt) {
      atomicAdd(&mesh[((size_t)ix * n + iy) * n + iz], wt);
      msum += (double)wt;
    });
  }
  const double m = block_sum_f64(msum, red, ATM_NT);
  if (threadIdx.x == 0 && m != 0.0) atomicAdd(&stats[2], m);
}

// ------------------------------------------------------------------------------------
// mass finalize: stats[0] = M (fixed-order sum of tile masses + atomic part), stats[1] = n^3 / M
// ------------------------------------------------------------------------------------
// fixed-order FP64 sum of the tile masses: each thread adds its slots in order (eight
// loads in flight per step), then wave butterflies and the wave totals
constexpr int MASS_NT = 1024;
__device__ __forceinline__ void mass_finalize_body(const Job& j, int ntiles) {
  __shared__ double red[MASS_NT / 64 + 1];
  constexpr int MU = 8;
  double s = 0.0;
  for (int i0 = threadIdx.x; i0 < ntiles; i0 += MASS_NT * MU) {
    double v[MU];
#pragma unroll
    for (int u = 0; u < MU; ++u) {
      const int i = i0 + u * MASS_NT;
      v[u] = i < ntiles ? j.tile_mass[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < MU; ++u) s += v[u];
  }
  const double m = block_sum_f64_waves<MASS_NT>(s, red) + j.stats[2];
  if (threadIdx.x == 0) {
    const double n = (double)j.a.n;
    j.stats[0] = m;
    j.stats[1] = n * n * n / m;
  }
  if (j.tile_sq != nullptr) {   // fused tiles: the sum of (rho - c)^2 in a fixed order -> aux[2]
    double q = 0.0;
    for (int i = threadIdx.x; i < ntiles; i += MASS_NT) q += j.tile_sq[i];
    const double tq = block_sum_f64_waves<MASS_NT>(q, red);
    if (threadIdx.x == 0) j.aux[2] = tq;
  }
}
__global__ void __launch_bounds__(MASS_NT) k_mass_finalize(Job j, int ntiles) { mass_finalize_body(j, ntiles); }


// ------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------
constexpr size_t LDS_MAX = 160 * 1024;
constexpr size_t LDS_TABLE_MAX = LDS_MAX - 16384;   // per-key tables of count / place in LDS
constexpr int CHUNK_MIN = 4096;          // smallest counting chunk (small catalogues)
constexpr int CHUNK_MAX = 1 << 22;       // largest (32-bit LDS counters above 65535)

static int largest_divisor_le(int n, int cap) {
  for (int d = cap < n ? cap : n; d >= 2; --d)
    if (n % d == 0) return d;
  return n;
}

void set_chunk(TileGeom* g, int64_t np, int64_t chunk) {
  if (chunk < CHUNK_MIN) chunk = CHUNK_MIN;
  if (chunk > CHUNK_MAX) chunk = CHUNK_MAX;
  g->chunk = (int)chunk;
  g->nblk = (int)((np + chunk - 1) / chunk);
  if (g->nblk < 1) g->nblk = 1;
}

size_t lds_table_keys() { return LDS_TABLE_MAX / 4; }

int make_tile_geom(int n, int64_t np, int stencil, const int* tile, TileGeom* g, bool* direct) {
  *direct = false;
  *g = TileGeom{};
  int ty = 0;
  if (tile) {
    if (tile[0] == 0) {
      *direct = true;
      return PK_OK;
    }
    ty = tile[1];
    if (tile[0] != 1 || (tile[2] != 0 && tile[2] != n))
      return set_error(PK_E_INVALID, "sorted deposit tiles are (1, ty, n) row bundles, got (%d, %d, %d)", tile[0],
                       tile[1], tile[2]);
    if (ty < 1 || ty > n || n % ty != 0) return set_error(PK_E_INVALID, "rows per tile %d must divide n=%d", ty, n);
    if (ty < 2 && n >= 2) return set_error(PK_E_INVALID, "rows per tile must be >= 2");
  } else {
    // largest divisor <= 32 whose rows fit the 64-bit fixed-point deposit (n <= 512: 32;
    // 768 and 1024: 16)
    ty = largest_divisor_le(n, 32);
    while (ty > 2 && ((size_t)ty * n * 8 > FIXED_LDS_MAX || n % ty != 0)) --ty;
  }
  g->ty = ty;
  g->nyt = n / ty;
  // exact for y * ty < 2^32 (y < n <= 4096): the multiplier exceeds 2^32 / ty by under 1
  g->ty_mul = (unsigned)((0xffffffffull + (unsigned long long)ty) / (unsigned long long)ty);
  g->ncls = stencil == PK_STENCIL_TSC ? 3 : 2;
  const long ntiles = (long)n * g->nyt;
  const long nkeys = ntiles * g->ncls + 1;
  const size_t rows_bytes = (size_t)ty * n * sizeof(float);
  if (rows_bytes > LDS_MAX - 8192) return set_error(PK_E_INVALID, "tile of %d rows x %d needs too much LDS", ty, n);
  g->ntiles = (int)ntiles;
  g->nkeys = (int)nkeys;
  g->fixed = (size_t)ty * n * 8 <= FIXED_LDS_MAX ? 1 : 0;
  // per-key tables of the count and place kernels: LDS when they fit, else the block's
  // row of the count matrix in HBM
  g->gtab = (size_t)nkeys * 4 > LDS_TABLE_MAX ? 1 : 0;
  // one counting / placing block per CU when the catalogue is large (a second, partial
  // round of blocks would cost a whole round), at least CHUNK_MIN particles per block
  int dev = 0;
  (void)hipGetDevice(&dev);
  const int64_t ncu = num_cus(dev);
  set_chunk(g, np, (np + ncu - 1) / ncu);
  if (g->nblk < 1) g->nblk = 1;
  if ((long)g->nkeys > (long)SCAN_BLOCK * SCAN_BLOCK || g->nblk > SCAN_BLOCK * SCAN_BLOCK)
    return set_error(PK_E_INVALID, "too many buckets");
  return PK_OK;
}

namespace {
int grid_for(int64_t items, int nt, int cap) {
  int64_t gsz = (items + nt - 1) / nt;
  if (gsz < 1) gsz = 1;
  if (gsz > cap) gsz = cap;
  return (int)gsz;
}

// raises a kernel's dynamic-LDS limit once per size (not again inside a graph capture)
template <typename K>
int allow_lds(K kernel, size_t bytes) {
  if (bytes <= 65536) return PK_OK;
  static std::mutex mu;
  static std::vector<std::pair<const void*, size_t>> done;
  const void* k = reinterpret_cast<const void*>(kernel);
  std::lock_guard<std::mutex> lk(mu);
  for (auto& d : done)
    if (d.first == k && d.second >= bytes) return PK_OK;
  int e = check_hip(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes), "LDS attribute");
  if (!e) done.emplace_back(k, bytes);
  return e;
}

size_t hist_lds(const TileGeom& g) {
  if (g.gtab) return 0;
  return g.chunk <= 65535 ? (size_t)((g.nkeys + 1) / 2) * 4 : (size_t)g.nkeys * 4;
}
}  // namespace

size_t carve_scatter(Job& j, bool direct, void* base) {
  size_t off = 0;
  char* b = static_cast<char*>(base);
  auto take = [&](size_t bytes) -> void* {
    void* p = b ? b + off : nullptr;
    off += align_up(bytes);
    return p;
  };
  const TileGeom& g = j.g;
  double* st = static_cast<double*>(take(PK_STATS_LEN * sizeof(double)));
  if (!j.stats) j.stats = st;
  if (!direct) {
    j.mat = static_cast<unsigned*>(take((size_t)g.nblk * g.nkeys * sizeof(unsigned)));
    j.totals = static_cast<int*>(take((size_t)g.nkeys * sizeof(int)));
    j.key_start = static_cast<int*>(take(((size_t)g.nkeys + 1) * sizeof(int)));
    j.bsum = static_cast<int*>(take((size_t)SCAN_BLOCK * sizeof(int)));
    j.wsum = static_cast<double*>(take((size_t)g.nblk * sizeof(double)));
    j.wmax = static_cast<double*>(take((size_t)g.nblk * sizeof(double)));
    j.aux = static_cast<double*>(take(AUX_LEN * sizeof(double)));
    j.koff = static_cast<unsigned*>(take((size_t)g.nkeys * sizeof(unsigned)));
    j.sorted = take((size_t)(j.a.np > 0 ? j.a.np : 1) * sizeof(float4));
    j.staged = take((size_t)(j.a.np > 0 ? j.a.np : 1) * sizeof(float4));
    j.tile_mass = static_cast<double*>(take((size_t)g.ntiles * sizeof(double)));
    j.tile_sq = static_cast<double*>(take((size_t)g.ntiles * sizeof(double)));
  }
  return off;
}

size_t scatter_workspace_bytes(const TileGeom& g, bool direct, int64_t np) {
  Job j{};
  j.g = g;
  j.a.np = np;
  return carve_scatter(j, direct, nullptr);
}

int mass_finalize(const Job& j, hipStream_t s) {
  StageTimer t(ST_MASS, s);
  hipLaunchKernelGGL(k_mass_finalize, dim3(1), dim3(MASS_NT), 0, s, j, j.tile_mass ? j.g.ntiles : 0);
  return check_hip(hipGetLastError(), "mass finalize launch");
}

int scatter_run(Job& j, bool direct, const float* mesh_in, float* mesh_out, void* ws, size_t ws_bytes,
                hipStream_t s) {
  const size_t need = carve_scatter(j, direct, ws);
  if (ws_bytes < need) return set_error(PK_E_WORKSPACE, "scatter workspace %zu < %zu bytes", ws_bytes, need);
  const ScatterArgs& a = j.a;
  const TileGeom& g = j.g;
  const int64_t ncell = (int64_t)a.n * a.n * a.n;
  if (direct) {
    {
      StageTimer t(ST_FILL, s);
      hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(PK_STATS_LEN), 0, s, j.stats, PK_STATS_LEN, 0.0);
      hipLaunchKernelGGL(k_copy_or_zero, dim3(grid_for(ncell, 256, 16384)), dim3(256), 0, s, mesh_in, mesh_out,
                         ncell);
      int e = check_hip(hipGetLastError(), "scatter fill launch");
      if (e) return e;
    }
    if (!mesh_out) return set_error(PK_E_INVALID, "direct mode needs a mesh");
    if (a.np > 0) {
      StageTimer t(ST_ATOMIC, s);
      hipLaunchKernelGGL(k_atomic_deposit, dim3(grid_for(a.np, ATM_NT, 65535)), dim3(ATM_NT), 0, s, a,
                         (const void*)nullptr, (const int*)nullptr, mesh_out, j.stats);
      int e = check_hip(hipGetLastError(), "atomic deposit launch");
      if (e) return e;
    }
    j.tile_mass = nullptr;
    j.tile_sq = nullptr;
    return mass_finalize(j, s);
  }
  {
    StageTimer t(ST_SORT_COUNT, s);
    const size_t lds = hist_lds(g);
    const auto kc = pick_aw<CountK>(a);
    int e = allow_lds(kc, lds);
    if (e) return e;
    hipLaunchKernelGGL(kc, dim3(g.nblk), dim3(CNT_NT), lds, s, j);
    e = check_hip(hipGetLastError(), "bucket count launch");
    if (e) return e;
  }
  {
    StageTimer t(ST_SCAN, s);
    hipLaunchKernelGGL(k_col_scan, dim3((g.nkeys + COL_KEYS - 1) / COL_KEYS), dim3(COL_NT), 0, s, j);
    const int nb2 = (g.nkeys + SCAN_BLOCK - 1) / SCAN_BLOCK;
    if (nb2 == 1) {
      hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(SCAN_NT), 0, s, j);
    } else {
      hipLaunchKernelGGL(k_scan_local, dim3(nb2), dim3(SCAN_NT), 0, s, j);
      hipLaunchKernelGGL(k_scan_add, dim3(nb2), dim3(SCAN_NT), 0, s, j);
    }
    int e = check_hip(hipGetLastError(), "scan launch");
    if (e) return e;
  }
  {
    StageTimer t(ST_SORT_PLACE, s);
    const size_t lds = place_lds(a.n);
    const auto kp = pick_aw<PlaceK>(a);
    int e = allow_lds(kp, lds);
    if (e) return e;
    hipLaunchKernelGGL(kp, dim3(g.nblk), dim3(CNT_NT), lds, s, j);
    e = check_hip(hipGetLastError(), "bucket place launch");
    if (e) return e;
  }
  {
    StageTimer t(ST_SORT_FINE, s);
    const auto kf = a.stencil == PK_STENCIL_CIC_REF ? k_fine_place<PK_STENCIL_CIC_REF>
                    : a.stencil == PK_STENCIL_CIC   ? k_fine_place<PK_STENCIL_CIC>
                                                    : k_fine_place<PK_STENCIL_TSC>;
    int e = allow_lds(kf, FINE_LDS);
    if (e) return e;
    hipLaunchKernelGGL(kf, dim3(fine_place_blocks(a.np)), dim3(FINE_NT), FINE_LDS, s, j);
    e = check_hip(hipGetLastError(), "fine place launch");
    if (e) return e;
  }
  if (!mesh_out) return PK_OK;   // the caller deposits (fused with the z pass)
  {
    StageTimer t(ST_DEPOSIT, s);
    const size_t lds = (size_t)g.ty * a.n * (g.fixed ? 8 : 4);
    int e;
    if (g.fixed) {
      e = allow_lds(k_rows_deposit<true>, lds);
      if (e) return e;
      hipLaunchKernelGGL(k_rows_deposit<true>, dim3(g.ntiles), dim3(DEP_NT), lds, s, a, g, (const int*)j.key_start,
                         (const void*)j.sorted, (const double*)j.aux, mesh_in, mesh_out, j.tile_mass);
    } else {
      e = allow_lds(k_rows_deposit<false>, lds);
      if (e) return e;
      hipLaunchKernelGGL(k_rows_deposit<false>, dim3(g.ntiles), dim3(DEP_NT), lds, s, a, g, (const int*)j.key_start,
                         (const void*)j.sorted, (const double*)j.aux, mesh_in, mesh_out, j.tile_mass);
    }
    e = check_hip(hipGetLastError(), "row deposit launch");
    if (e) return e;
  }
  if (a.stencil == PK_STENCIL_CIC_REF && a.np > 0) {
    StageTimer t(ST_ATOMIC, s);
    hipLaunchKernelGGL(k_atomic_deposit, dim3(64), dim3(ATM_NT), 0, s, a, (const void*)j.sorted,
                       (const int*)(j.key_start + g.nkeys - 1), mesh_out, j.stats);
    int e = check_hip(hipGetLastError(), "irregular deposit launch");
    if (e) return e;
  }
  j.tile_sq = nullptr;   // mesh tiles: the z pass sums the squares
  return mass_finalize(j, s);
}

int scatter_sort_jobs(const Job* jobs, const BatchDims& d, hipStream_t s) {
  const int nj = d.njobs;
  {
    StageTimer t(ST_SORT_COUNT, s);
    const size_t lds = (size_t)d.lds_keys * 4;   // any job may count with 32-bit counters
    int e = allow_lds(k_bucket_count_jobs, lds);
    if (e) return e;
    hipLaunchKernelGGL(k_bucket_count_jobs, dim3(d.max_nblk, nj), dim3(CNT_NT), lds, s, jobs);
    e = check_hip(hipGetLastError(), "batch bucket count launch");
    if (e) return e;
  }
  {
    StageTimer t(ST_SCAN, s);
    const int nb2 = (d.max_nkeys + SCAN_BLOCK - 1) / SCAN_BLOCK;
    hipLaunchKernelGGL(k_col_scan_jobs, dim3((d.max_nkeys + COL_KEYS - 1) / COL_KEYS, nj), dim3(COL_NT), 0, s, jobs);
    hipLaunchKernelGGL(k_scan_local_jobs, dim3(nb2, nj), dim3(SCAN_NT), 0, s, jobs);
    hipLaunchKernelGGL(k_scan_add_jobs, dim3(nb2, nj), dim3(SCAN_NT), 0, s, jobs);
    int e = check_hip(hipGetLastError(), "batch scan launch");
    if (e) return e;
  }
  {
    StageTimer t(ST_SORT_PLACE, s);
    const size_t lds = place_lds(d.n);
    int e = allow_lds(k_bucket_place_jobs, lds);
    if (e) return e;
    hipLaunchKernelGGL(k_bucket_place_jobs, dim3(d.max_nblk, nj), dim3(CNT_NT), lds, s, jobs);
    e = check_hip(hipGetLastError(), "batch bucket place launch");
    if (e) return e;
  }
  StageTimer t(ST_SORT_FINE, s);
  int e = allow_lds(k_fine_place_jobs, FINE_LDS);
  if (e) return e;
  hipLaunchKernelGGL(k_fine_place_jobs, dim3(d.max_fblk, nj), dim3(FINE_NT), FINE_LDS, s, jobs);
  return check_hip(hipGetLastError(), "batch fine place launch");
}


}  // namespace pk
