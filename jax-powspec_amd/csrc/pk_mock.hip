// Seeded lognormal mock generator: Gaussian half spectrum (src/gauss_field.py:5-80),
// Poisson population of a density field (src/gauss_field.py:83-110,
// src/populate_field.py:4-29) with the lognormal step of tests/create_lognormal.py:50-51
// fused in, and an optional Zel'dovich redshift-space shift along z.
//
// Population passes:
//   sum   : FP64 per-block sums of rho = exp(bias * field) (or field) -> fixed-order total
//   count : Poisson counts per cell (one byte each, saturated) -> per-block totals ->
//           exclusive scan (block offsets)
//   mark  : block-local scans of the stored counts; each particle slot receives its cell
//   emit  : one thread per particle: draws, offset, shift, wrap, the (x, y, z) record
// A block owns POP_CELLS consecutive cells in C order, visited in rounds of 2 POP_NT
// consecutive cells (a thread takes a pair, whose two Poisson draws share one Philox
// block), so particles come out in cell order and every offset is a prefix sum.  Emitting
// per particle rather than per cell keeps the random-number and FP64 work free of the
// divergence of sparse cells (most cells hold no particle).
#include "pk_common.h"
#include "pk_internal.h"
#include "pk_mock.h"

namespace pk {

namespace {
constexpr int AMP_NT = 256;
constexpr int SPEC_NT = 256;
constexpr int POP_NT = 256;
constexpr int POP_ROUNDS = 8;
constexpr int POP_CELLS = 2 * POP_NT * POP_ROUNDS;
constexpr int TOP_NT = 1024;
constexpr int MISC_LEN = 4;   // [0] rate scale, [1] total (as int64 bits), [2] sum rho

int64_t pop_blocks(int n) {
  const int64_t nc = (int64_t)n * n * n;
  return (nc + POP_CELLS - 1) / POP_CELLS;
}
}  // namespace

size_t gauss_table_len(int n) { return (size_t)3 * (size_t)(n / 2) * (size_t)(n / 2) + 1; }

size_t populate_ws_bytes(int n) {
  const size_t nb = (size_t)pop_blocks(n);
  const size_t npair = ((size_t)n * n * n + 1) / 2;
  return align_up(nb * sizeof(double)) + align_up(nb * sizeof(long long)) + align_up(MISC_LEN * sizeof(double)) +
         align_up(npair * sizeof(uint16_t));
}

// ------------------------------------------------------------------------------------
// amplitude table: sqrt(P(k_F sqrt(m)) (n^2/L)^3), the bisection and linear
// interpolation of src/gauss_field.py:39-47 (end segments extrapolate)
// ------------------------------------------------------------------------------------
__global__ void __launch_bounds__(AMP_NT) k_amp_table(const double* __restrict__ kf, const double* __restrict__ pkf,
                                                      int nk, double kpref, double units, int mmax,
                                                      float* __restrict__ amp) {
  const int m = blockIdx.x * AMP_NT + threadIdx.x;
  if (m > mmax) return;
  const double kmod = sqrt((double)m) * kpref;
  int lo = 0, hi = nk - 1;
  while (hi - lo > 1) {
    const int l = (lo + hi) >> 1;
    if (kf[l] < kmod) lo = l;
    else hi = l;
  }
  const double P = (pkf[hi] - pkf[lo]) / (kf[hi] - kf[lo]) * (kmod - kf[lo]) + pkf[lo];
  amp[m] = (float)sqrt(P * units);
}

int gauss_table(int n, double box, const double* kf, const double* pkf, int nk, float* amp, hipStream_t s) {
  const int mmax = (int)gauss_table_len(n) - 1;
  const double kpref = 2.0 * M_PI / box;
  const double u = (double)n * (double)n / box;
  const double units = u * u * u;
  hipLaunchKernelGGL(k_amp_table, dim3((mmax + AMP_NT) / AMP_NT), dim3(AMP_NT), 0, s, kf, pkf, nk, kpref, units, mmax,
                     amp);
  return check_hip(hipGetLastError(), "amplitude table launch");
}

// the stored half spectrum itself, [kz][kx][ky]
__global__ void __launch_bounds__(SPEC_NT) k_gauss_spectrum(FieldGen g, float2* __restrict__ spec, int64_t total) {
  const int n = g.n;
  const int64_t nn = (int64_t)n * n;
  for (int64_t i = (int64_t)blockIdx.x * SPEC_NT + threadIdx.x; i < total; i += (int64_t)gridDim.x * SPEC_NT) {
    const int kz = (int)(i / nn);
    const int r = (int)(i - (int64_t)kz * nn);
    spec[i] = gauss_mode(g, r / n, r % n, kz);
  }
}

int gauss_spectrum(const FieldGen& g, float2* spec, hipStream_t s) {
  const int64_t total = (int64_t)(g.n / 2 + 1) * g.n * g.n;
  int64_t blocks = (total + SPEC_NT - 1) / SPEC_NT;
  int dev = 0;
  (void)hipGetDevice(&dev);
  const int64_t cap = (int64_t)num_cus(dev) * 32;
  if (blocks > cap) blocks = cap;
  hipLaunchKernelGGL(k_gauss_spectrum, dim3((unsigned)blocks), dim3(SPEC_NT), 0, s, g, spec, total);
  return check_hip(hipGetLastError(), "Gaussian spectrum launch");
}

// ------------------------------------------------------------------------------------
// population
// ------------------------------------------------------------------------------------
struct PopWs {
  double* rsum;      // [nblk]
  long long* boff;   // [nblk] counts, then exclusive offsets
  double* misc;      // [MISC_LEN]
  uint16_t* cnt;     // [ncell / 2] counts of cell pairs (two saturated bytes)
};

static PopWs carve_pop(void* ws, int n) {
  const size_t nb = (size_t)pop_blocks(n);
  char* p = static_cast<char*>(ws);
  PopWs w;
  w.rsum = reinterpret_cast<double*>(p);
  p += align_up(nb * sizeof(double));
  w.boff = reinterpret_cast<long long*>(p);
  p += align_up(nb * sizeof(long long));
  w.misc = reinterpret_cast<double*>(p);
  p += align_up(MISC_LEN * sizeof(double));
  w.cnt = reinterpret_cast<uint16_t*>(p);
  return w;
}

struct PopK {
  const float* field;
  const float* psi;
  int64_t ncell;
  int n;
  int nblk;
  double bias;
  double mean_obj;   // cell volume * density
  double box, H, growth;
  uint32_t k0, k1;
  double* rsum;
  long long* boff;
  double* misc;
  uint16_t* cnt;
  long long* total;  // caller's device int64 (nullable)
  float* pos;
  long long capacity;
};

__device__ __forceinline__ double cell_rho(float v, double bias) {
  return bias != 0.0 ? exp(bias * (double)v) : (double)v;
}

// first cell of round k of this thread: a block owns POP_CELLS consecutive cells in C
// order, visited in POP_ROUNDS rounds of 2 POP_NT cells; thread t takes the pair
// (2t, 2t + 1) of each round
__device__ __forceinline__ int64_t pop_pair(int k) {
  return (int64_t)blockIdx.x * POP_CELLS + (int64_t)k * (2 * POP_NT) + 2 * (int64_t)threadIdx.x;
}

__device__ __forceinline__ long long block_sum_i64(long long v, long long* red) {
  constexpr int NW = POP_NT / 64;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  long long t = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) t += red[i];
  return t;
}

// exclusive prefix of v over the block (int64) and the block total; sh holds NT / 64
// entries and is free again only after the caller's next barrier
template <int NT>
__device__ __forceinline__ long long block_scan_i64(long long v, long long* sh, long long& total) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  long long incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const long long t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  if (lane == 63) sh[w] = incl;
  __syncthreads();
  long long before = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NT / 64; ++i) {
    const long long x = sh[i];
    if (i < w) before += x;
    tot += x;
  }
  total = tot;
  return before + incl - v;
}

__global__ void __launch_bounds__(POP_NT) k_pop_sum(PopK a) {
  __shared__ double red[POP_NT / 64 + 1];
  float v[2 * POP_ROUNDS];
#pragma unroll
  for (int k = 0; k < POP_ROUNDS; ++k) {
    const int64_t c = pop_pair(k);
    v[2 * k] = c < a.ncell ? a.field[c] : 0.0f;
    v[2 * k + 1] = c + 1 < a.ncell ? a.field[c + 1] : 0.0f;
  }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < POP_ROUNDS; ++k) {
    const int64_t c = pop_pair(k);
    if (c < a.ncell) s += cell_rho(v[2 * k], a.bias);
    if (c + 1 < a.ncell) s += cell_rho(v[2 * k + 1], a.bias);
  }
  const double t = block_sum_f64_waves<POP_NT>(s, red);
  if (threadIdx.x == 0) a.rsum[blockIdx.x] = t;
}

// fixed-order total of the block sums -> rate scale = cell volume * density / mean(rho)
__global__ void __launch_bounds__(TOP_NT) k_pop_scale(PopK a) {
  __shared__ double red[TOP_NT / 64 + 1];
  const int per = (a.nblk + TOP_NT - 1) / TOP_NT;
  const int b0 = threadIdx.x * per;
  double s = 0.0;
  for (int i = 0; i < per; ++i)
    if (b0 + i < a.nblk) s += a.rsum[b0 + i];
  const double t = block_sum_f64_waves<TOP_NT>(s, red);
  if (threadIdx.x == 0) {
    a.misc[0] = a.mean_obj / (t / (double)a.ncell);
    a.misc[2] = t;
  }
}

// counts of the pair (c, c + 1); cells beyond the mesh count 0
__device__ __forceinline__ void pair_counts(const PopK& a, int64_t c, double scale, int& n0, int& n1) {
  const double l0 = c < a.ncell ? cell_rho(a.field[c], a.bias) * scale : 0.0;
  const double l1 = c + 1 < a.ncell ? cell_rho(a.field[c + 1], a.bias) * scale : 0.0;
  poisson_pair(l0, l1, (uint64_t)c, a.k0, a.k1, n0, n1);
}

__global__ void __launch_bounds__(POP_NT) k_pop_count(PopK a) {
  __shared__ long long red[POP_NT / 64];
  const double scale = a.misc[0];
  long long s = 0;
#pragma unroll 1
  for (int k = 0; k < POP_ROUNDS; ++k) {
    const int64_t c = pop_pair(k);
    int n0, n1;
    pair_counts(a, c, scale, n0, n1);
    s += (long long)n0 + n1;
    // stored saturated: 255 means "recount"
    if (c < a.ncell) a.cnt[c >> 1] = (uint16_t)(min(n0, 255) | (min(n1, 255) << 8));
  }
  const long long t = block_sum_i64(s, red);
  if (threadIdx.x == 0) a.boff[blockIdx.x] = t;
}

// block totals -> exclusive block offsets (in place); grand total to misc[1] and *total
__global__ void __launch_bounds__(TOP_NT) k_pop_scan(PopK a) {
  __shared__ long long sh[TOP_NT / 64];
  const int per = (a.nblk + TOP_NT - 1) / TOP_NT;
  const int b0 = threadIdx.x * per;
  long long s = 0;
  for (int i = 0; i < per; ++i)
    if (b0 + i < a.nblk) s += a.boff[b0 + i];
  long long total;
  long long run = block_scan_i64<TOP_NT>(s, sh, total);
  for (int i = 0; i < per; ++i)
    if (b0 + i < a.nblk) {
      const long long c = a.boff[b0 + i];
      a.boff[b0 + i] = run;
      run += c;
    }
  if (threadIdx.x == 0) {
    reinterpret_cast<long long*>(a.misc)[1] = total;
    if (a.total) *a.total = total;
  }
}

// particle slots -> cell ids: the first word of each output record receives its cell
// (as uint32), which k_pop_emit turns into the position
__global__ void __launch_bounds__(POP_NT) k_pop_mark(PopK a) {
  __shared__ long long sh[2][POP_NT / 64];
  long long base = a.boff[blockIdx.x];
  uint32_t* rec = reinterpret_cast<uint32_t*>(a.pos);
#pragma unroll 1
  for (int k = 0; k < POP_ROUNDS; ++k) {
    const int64_t c = pop_pair(k);
    int n0 = 0, n1 = 0;
    if (c < a.ncell) {
      const uint16_t v = a.cnt[c >> 1];
      n0 = v & 255;
      n1 = v >> 8;
      if (n0 == 255 || n1 == 255) pair_counts(a, c, a.misc[0], n0, n1);
    }
    long long rtot;
    // two scan buffers: round k + 1 writes the other one, so one barrier per round
    const long long p = base + block_scan_i64<POP_NT>((long long)n0 + n1, sh[k & 1], rtot);
    base += rtot;
    // the wave's particles occupy the contiguous slots [p(lane 0), p(lane 63) + count(63));
    // they are written lane by lane (coalesced, no divergence): the owner of a slot is the
    // last lane whose first slot is not above it (lanes without particles share the next
    // lane's offset, and an owner's successors start past its slots)
    const int lane = threadIdx.x & 63;
    const long long w0 = __shfl(p, 0, 64);
    const int off = (int)(p - w0);
    const int T = (int)(__shfl(p + n0 + n1, 63, 64) - w0);
    const int64_t pair0 = c - 2 * (int64_t)lane;        // the wave's first cell
    for (int s0 = 0; s0 < T; s0 += 64) {                // wave-uniform trip count
      const int sidx = s0 + lane;
      int lo = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1) {        // largest lane L with off(L) <= sidx
        const int probe = __shfl(off, lo + step, 64);
        if (probe <= sidx) lo += step;
      }
      const int o = __shfl(off, lo, 64);
      const int first = __shfl(n0, lo, 64);
      if (sidx < T) {
        const long long slot = w0 + sidx;
        if (slot < a.capacity) {
          const int64_t cell = pair0 + 2 * (int64_t)lo + (sidx - o < first ? 0 : 1);
          rec[3 * slot] = (uint32_t)cell;
        }
      }
    }
  }
}

// periodic wrap into [0, L) in FP64, rounded to float32 (a value that rounds up to L
// becomes 0); values already inside skip the division
__device__ __forceinline__ float wrap_box(double r, double box) {
  double u = r;
  if (!(u >= 0.0 && u < box)) {
    u = r - box * floor(r / box);
    if (u < 0.0) u += box;
    if (u >= box) u -= box;
  }
  const float f = (float)u;
  return f >= (float)box ? 0.0f : f;
}

// one thread per particle: cell centre + triangular offset (+ redshift-space shift),
// wrapped into [0, L)
__global__ void __launch_bounds__(POP_NT) k_pop_emit(PopK a) {
  const long long p = (long long)blockIdx.x * POP_NT + threadIdx.x;
  const long long total = reinterpret_cast<const long long*>(a.misc)[1];
  if (p >= a.capacity || p >= total) return;
  const uint32_t c = reinterpret_cast<const uint32_t*>(a.pos)[3 * p];
  const uint32_t n = (uint32_t)a.n;
  const uint32_t iz = c % n, cy = c / n;
  const uint32_t iy = cy % n, ix = cy / n;
  const Philox4 w = rng_block((uint64_t)p, 0, TAG_DISP, a.k0, a.k1);
  const double shift = (a.psi != nullptr && a.growth != 0.0) ? a.growth * (double)a.psi[c] : 0.0;
  double r[3] = {((double)ix + 0.5 + tri_offset(w.x)) * a.H, ((double)iy + 0.5 + tri_offset(w.y)) * a.H,
                 ((double)iz + 0.5 + tri_offset(w.z)) * a.H + shift};
  float f[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) f[d] = wrap_box(r[d], a.box);
  a.pos[3 * p] = f[0];
  a.pos[3 * p + 1] = f[1];
  a.pos[3 * p + 2] = f[2];
}

namespace {
PopK make_pop(const float* field, const float* psi, int n, double box, double density, double bias, double growth,
              uint64_t seed, void* ws, long long* total, float* pos, long long capacity) {
  PopK a{};
  a.field = field;
  a.psi = psi;
  a.n = n;
  a.ncell = (int64_t)n * n * n;
  a.nblk = (int)pop_blocks(n);
  a.bias = bias;
  a.box = box;
  a.H = box / (double)n;
  a.mean_obj = a.H * a.H * a.H * density;
  a.growth = growth;
  a.k0 = (uint32_t)seed;
  a.k1 = (uint32_t)(seed >> 32);
  const PopWs w = carve_pop(ws, n);
  a.rsum = w.rsum;
  a.boff = w.boff;
  a.misc = w.misc;
  a.cnt = w.cnt;
  a.total = total;
  a.pos = pos;
  a.capacity = capacity;
  return a;
}
}  // namespace

int populate_count(const float* field, int n, double box, double density, double bias, uint64_t seed, void* ws,
                   int64_t* total, hipStream_t s) {
  const PopK a = make_pop(field, nullptr, n, box, density, bias, 0.0, seed, ws, reinterpret_cast<long long*>(total),
                          nullptr, 0);
  hipLaunchKernelGGL(k_pop_sum, dim3(a.nblk), dim3(POP_NT), 0, s, a);
  hipLaunchKernelGGL(k_pop_scale, dim3(1), dim3(TOP_NT), 0, s, a);
  hipLaunchKernelGGL(k_pop_count, dim3(a.nblk), dim3(POP_NT), 0, s, a);
  hipLaunchKernelGGL(k_pop_scan, dim3(1), dim3(TOP_NT), 0, s, a);
  return check_hip(hipGetLastError(), "population count launch");
}

int populate_place(const float* field, const float* psi, double growth, int n, double box, double bias, uint64_t seed,
                   void* ws, float* pos, int64_t capacity, hipStream_t s) {
  if (capacity == 0) return PK_OK;
  const PopK a = make_pop(field, psi, n, box, 0.0, bias, growth, seed, ws, nullptr, pos, (long long)capacity);
  hipLaunchKernelGGL(k_pop_mark, dim3(a.nblk), dim3(POP_NT), 0, s, a);
  const long long g = (capacity + POP_NT - 1) / POP_NT;
  hipLaunchKernelGGL(k_pop_emit, dim3((unsigned)g), dim3(POP_NT), 0, s, a);
  return check_hip(hipGetLastError(), "population place launch");
}

}  // namespace pk
