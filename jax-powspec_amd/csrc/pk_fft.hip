// 3D real-to-complex FFT passes and the fused x-pass + deconvolution + Legendre binning.
//
// Spectrum layout: float2 [n/2+1][n][n] indexed [kz][kx][ky] (ky fastest), so
//   z pass : real rows along z (contiguous in the mesh) -> R2C via a length-n/2 complex
//            FFT + split; writes [kz][x][y] with 16+ consecutive y per 128 B segment;
//   y pass : contiguous rows along y, in place;
//   x pass : columns along x at stride n, tiles of CX consecutive ky; fused with the
//            binning so the final spectrum is never written back (pk_powspec_mesh),
//            or written back in place (pk_r2c).
// Every pass is a persistent grid (occupancy-sized) looping over tiles; each tile is
// staged in LDS and transformed with the Stockham stages of pk_common.h.
//
// Reference semantics (src/correlations.py:7-56, :60-117):
//   delta_k = rfftn(delta) (:22); correction prod_a (x_a / sin x_a)^p, x_a = pi k_a / n
//   (:20-21, :32, :39); P2 = |delta_k|^2 (:40); mu = kz / k, LOS = z (:36-38);
//   l = 0, 2, 4 weights (:45-47); every stored half-spectrum entry counts once (:48).
#include "pk_common.h"
#include "pk_internal.h"
#include "pk_mock.h"

namespace pk {

template <int N> struct FFTCfg {
  static constexpr int NC = N / 2;
  // z pass: CZ real rows (consecutive y) per tile
  static constexpr int CZ = cmin(pow2_div(N), cmin(64, cmax(8, pow2_floor(4096 / NC))));
  static constexpr int NTZ = cmin(1024, cmax(64, (CZ * NC / 16) / 64 * 64));
  // y pass: CY rows per tile
  static constexpr int CY = cmin(pow2_div(N), cmin(64, cmax(1, pow2_floor(4096 / N))));
  static constexpr int NTY = cmin(1024, cmax(64, (CY * N / 16) / 64 * 64));
  // x pass: CX columns (consecutive ky) per tile
  static constexpr int CX = cmin(pow2_div(N), cmin(64, cmax(N >= 1024 ? 8 : 16, pow2_floor(4096 / N))));
  static constexpr int NTX = cmin(1024, cmax(64, (CX * N / 16) / 64 * 64));
  // z passes: waves per SIMD their LDS allows (tile, twiddles, FP64 reduction; capped at
  // 4), to which their register budget is held
  static constexpr int WPEZ = cmax(
      1, cmin(4, (int)((160 * 1024) / ((size_t)CZ * (NC + 1) * 8 + (size_t)N * 8 + (size_t)NTZ * 8 + 1024)) *
                     (NTZ / 64) / 4));
};

// ------------------------------------------------------------------------------------
// z pass: real rows -> half spectrum rows, written transposed to [kz][x][y]
// ------------------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(FFTCfg<N>::NTZ) __attribute__((amdgpu_waves_per_eu(FFTCfg<N>::WPEZ)))
k_zpass(const float* __restrict__ mesh, float2* __restrict__ spec, const float2* __restrict__ twg,
        const double* __restrict__ stats, double* __restrict__ dc_part, double* __restrict__ sq_part, int kz_lim) {
  using Cf = FFTCfg<N>;
  constexpr int NC = Cf::NC, C = Cf::CZ, NT = Cf::NTZ, CS = NC + 1;
  constexpr int YT = N / C;
  constexpr int NTILES = N * YT;
  __shared__ float2 buf[C * CS];
  __shared__ float2 tw[N];
  __shared__ double red[NT];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += NT) tw[i] = twg[i];
  double dc = 0.0;  // FP64 sum of the transformed values = the k = 0 mode
  double sq = 0.0;  // and of their squares (Parseval bound of the binned power)
  float s = 1.0f, b = 0.0f;
  if (stats != nullptr) {  // contrast delta = rho / mean - 1 (tests/lognormal.py:79-80)
    s = (float)stats[1];
    b = -1.0f;
  }
  for (int tile = blockIdx.x; tile < NTILES; tile += gridDim.x) {
    const int x = tile / YT, yt = tile % YT;
    const float2* src = reinterpret_cast<const float2*>(mesh + ((size_t)x * N + (size_t)yt * C) * N);
    {
      // all loads in flight first; the FP64 k = 0 sum keeps its per-thread order
      constexpr int TOT = C * NC, PER = (TOT + NT - 1) / NT;
      float2 v[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int i = tid + k * NT;
        if (TOT % NT == 0 || i < TOT) v[k] = src[i];
      }
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int i = tid + k * NT;
        if (TOT % NT == 0 || i < TOT) {
          v[k].x = v[k].x * s + b;
          v[k].y = v[k].y * s + b;
          dc += (double)v[k].x + (double)v[k].y;
          sq += (double)v[k].x * v[k].x + (double)v[k].y * v[k].y;
          buf[(i / NC) * CS + (i % NC)] = v[k];
        }
      }
    }
    __syncthreads();
    lds_fft<NC, C, 1, CS, NT, true, 2>(buf, tw);
    // split: X[k] = Fe + e^{-2 pi i k / N} Fo, Fe = (Z_k + Z*_{NC-k}) / 2, Fo = -i (Z_k - Z*_{NC-k}) / 2
    // output planes kz <= kz_lim only (callers that read no higher plane pass less than n/2)
    for (int i = tid; i < (min(kz_lim, NC) + 1) * C; i += NT) {
      const int c = i % C, k = i / C;
      const float2 zk = buf[c * CS + (k == NC ? 0 : k)];
      const float2 zr = buf[c * CS + (k == 0 ? 0 : NC - k)];
      const float2 zc = make_float2(zr.x, -zr.y);
      const float2 fe = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y + zc.y));
      const float2 dd = csub(zk, zc);
      const float2 fo = make_float2(0.5f * dd.y, -0.5f * dd.x);
      const float2 X = cadd(fe, cmul(tw[k], fo));
      spec[((size_t)k * N + x) * N + (size_t)yt * C + c] = X;
    }
    __syncthreads();
  }
  if (dc_part != nullptr) {  // fixed-order tree: the k = 0 mode in FP64 (sum of all values)
    const double t = block_sum_f64(dc, red, NT);
    if (tid == 0) dc_part[blockIdx.x] = t;
  }
  if (sq_part != nullptr) {
    const double t = block_sum_f64(sq, red, NT);
    if (tid == 0) sq_part[blockIdx.x] = t;
  }
}

// ------------------------------------------------------------------------------------
// y pass: contiguous rows of N complex, in place
// ------------------------------------------------------------------------------------
// CONJ: conj on load and store, i.e. the unnormalised inverse transform (e^{+i})
// ntiles < 0: every tile; otherwise the first ntiles (rows [kz][x] of the lowest planes)
template <int N, bool CONJ = false>
__device__ __forceinline__ void ypass_body(float2* __restrict__ spec, const float2* __restrict__ twg, int bx, int nbx,
                                           int ntiles = -1) {
  using Cf = FFTCfg<N>;
  // rows are skewed in LDS (point p at p + p / 8), which keeps the butterflies' strided
  // writes off each other's banks; twiddles from a quarter table
  constexpr int SK = 3;
  constexpr int C = Cf::CY, NT = Cf::NTY, CS = N + (SK > 0 ? (N >> SK) : 0) + 1;
  const int NTILES = ntiles < 0 ? (N / 2 + 1) * N / C : ntiles;
  __shared__ float2 buf[C * CS];
  __shared__ float2 twh[N / 2];
  const TwHalf<N> tw{twh};
  constexpr int PER = (C * N + NT - 1) / NT;
  const int tid = threadIdx.x;
  for (int i = tid; i < N / 2; i += NT) twh[i] = twg[i];
  auto at = [](int p, int c) { return c * CS + p + (SK > 0 ? (p >> SK) : 0); };
  // the next tile's rows are loaded into registers while this tile is transformed
  float2 pre[PER];
  auto load = [&](int tile) {
    const float2* row = spec + (size_t)tile * C * N;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = tid + k * NT;
      if (i < C * N) {
        pre[k] = row[i];
        if constexpr (CONJ) pre[k].y = -pre[k].y;
      }
    }
  };
  // Work item t -> tile.  Whole passes take the rows x-plane major from the highest x down:
  // the z pass writes the rows x-plane by x-plane upwards, so the first rows read here are
  // the last it wrote.
  auto tile_of = [&](int t) -> int {
    if (ntiles >= 0) return t;
    constexpr int KZN = N / 2 + 1, XG = N / C;
    const int u = NTILES - 1 - t;
    return (u % KZN) * XG + u / KZN;
  };
  if (bx < NTILES) load(tile_of(bx));
  for (int item = bx; item < NTILES; item += nbx) {
    const int tile = tile_of(item);
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = tid + k * NT;
      if (i < C * N) buf[at(i % N, i / N)] = pre[k];
    }
    __syncthreads();
    if (item + nbx < NTILES) load(tile_of(item + nbx));
    lds_fft<N, C, 1, CS, NT, false, 1, SK>(buf, tw);
    // Whole passes store the first half of their rows (the high x planes) non-temporally:
    // the last-level cache cannot hold the pass's output, and those rows would be evicted
    // before the x pass reads them, so they go to memory without displacing the second
    // half (512^3: y pass 206 -> 185 us, the next passes faster too; a quarter, three
    // quarters or all of the rows measured slower)
    float2* row = spec + (size_t)tile * C * N;
    const bool streamed = ntiles < 0 && item < NTILES / 2;   // block-uniform
    typedef float f2v __attribute__((ext_vector_type(2)));
    for (int i = tid; i < C * N; i += NT) {
      float2 v = buf[at(i % N, i / N)];
      if constexpr (CONJ) v.y = -v.y;
      if (streamed) {
        const f2v w = {v.x, v.y};
        __builtin_nontemporal_store(w, reinterpret_cast<f2v*>(&row[i]));
      } else {
        row[i] = v;
      }
    }
    __syncthreads();
  }
}

// The mass finalize of a fused estimate, done by one block of NT threads: stats[0] = M (the
// tiles' exact masses summed in a fixed order, plus stats[2]), stats[1] = n^3 / M, and
// aux[2] = the tiles' sums of (rho - c)^2 in a fixed order (the binning's fixed-point
// bound).  U loads per array in flight per thread.
template <int NT>
__device__ __forceinline__ void mass_finalize_block(const Job& j) {
  __shared__ double red[NT / 64 + 1];
  constexpr int U = 16;
  const int nt = j.g.ntiles;
  double m = 0.0, q = 0.0;
  for (int i0 = threadIdx.x; i0 < nt; i0 += NT * U) {   // block-uniform trip count
    double vm[U], vq[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = min(i0 + u * NT, nt - 1);
      vm[u] = j.tile_mass[i];
      vq[u] = j.tile_sq[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (i0 + u * NT < nt) {
        m += vm[u];
        q += vq[u];
      }
    }
  }
  const double tm = block_sum_f64_waves<NT>(m, red) + j.stats[2];
  const double tq = block_sum_f64_waves<NT>(q, red);
  if (threadIdx.x == 0) {
    const double n = (double)j.a.n;
    j.stats[0] = tm;
    j.stats[1] = n * n * n / tm;
    j.aux[2] = tq;
  }
}

template <int N>
__global__ void __launch_bounds__(FFTCfg<N>::NTY) k_ypass(float2* __restrict__ spec, const float2* __restrict__ twg) {
  ypass_body<N>(spec, twg, blockIdx.x, gridDim.x);
}

// y pass of a fused estimate whose last block first does the mass finalize (the last block
// has at most as many tiles as any other)
template <int N>
__global__ void __launch_bounds__(FFTCfg<N>::NTY) k_ypass_mass(Job j, const float2* __restrict__ twg) {
  if (blockIdx.x == gridDim.x - 1) mass_finalize_block<FFTCfg<N>::NTY>(j);
  ypass_body<N>(j.spec, twg, blockIdx.x, gridDim.x);
}

template <int N>
__global__ void __launch_bounds__(FFTCfg<N>::NTY) k_ypass_lim(float2* __restrict__ spec,
                                                              const float2* __restrict__ twg, int kz_lim) {
  ypass_body<N>(spec, twg, blockIdx.x, gridDim.x, (kz_lim + 1) * N / FFTCfg<N>::CY);
}

template <int N>
__global__ void __launch_bounds__(FFTCfg<N>::NTY) k_ypass_jobs(const Job* __restrict__ jobs,
                                                               const float2* __restrict__ twg) {
  const Job& j = jobs[blockIdx.y];
  if (!j.active) return;
  if (blockIdx.x == gridDim.x - 1) mass_finalize_block<FFTCfg<N>::NTY>(j);   // fused jobs
  ypass_body<N>(j.spec, twg, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------------------------------------
// x pass (plain): columns along x, written back in place
// ------------------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(FFTCfg<N>::NTX)
k_xpass(float2* __restrict__ spec, const float2* __restrict__ twg, int kz_lim) {
  using Cf = FFTCfg<N>;
  constexpr int C = Cf::CX, NT = Cf::NTX, PS = C + 1;
  constexpr int YT = N / C;
  constexpr int NTILES = (N / 2 + 1) * YT;
  __shared__ float2 buf[N * PS];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += NT) tw[i] = twg[i];
  const int ntiles = min(NTILES, (kz_lim + 1) * YT);
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int kz = tile / YT, yt = tile % YT;
    float2* base = spec + (size_t)kz * N * N + (size_t)yt * C;
    stage_tile<N, C, PS, NT>(buf, base, N);
    __syncthreads();
    lds_fft<N, C, PS, 1, NT, true, 1>(buf, tw);
    for (int i = tid; i < N * C; i += NT) base[(size_t)(i / C) * N + (i % C)] = buf[(i / C) * PS + (i % C)];
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------
// bin lookup: slot of m = kx^2+ky^2+kz^2 under integer thresholds (see pk_api.h)
// ------------------------------------------------------------------------------------
template <class JT>
__device__ __forceinline__ int find_bin(int m, const int* __restrict__ thr, const JT* __restrict__ jst, int nb) {
  int j = (int)sqrtf((float)m);
  if (j * j > m) --j;
  else if ((j + 1) * (j + 1) <= m) ++j;
  int b = (int)jst[j];
  while (b + 1 < nb && thr[b + 1] <= m) ++b;
  if (b < 0 || m >= thr[nb]) return -1;
  return b;
}

constexpr int HB_LDS = 256;   // bins held in LDS; above this the per-block slice lives in HBM
constexpr int NACC = BIN_PART_LEN;   // S0, S2, S4 and the binned mode count per block and bin

// Exponent e of the bins' 64-bit fixed point (a sum is an integer times 2^-e).  Every entry
// the x pass bins is a stored mode's |X_k|^2 times window factors of at most wmax^3 (wmax:
// the factor at |k| = n/2, the largest), and the stored modes' |X_k|^2 add up to at most
// n^3 sum(x^2) over the real field x that was transformed (Parseval).  So 2 wmax^3 n^3
// sum(x^2) (the 2 covers float rounding) bounds every bin's sums and their total, and e puts
// that bound at or below 2^62.  The partial sums of x^2 are added in a fixed order, so every
// block (and the finalize) gets the same e, and integer bins make the sums independent of the
// order of the atomics.  Every thread must call it (two barriers).
// finite (optional): false when the field holds a NaN or an infinity, whose power the
// finalize then reports as NaN (as FP64 sums of such a field would be).
template <int NT>
__device__ __forceinline__ int bin_fixed_exp(const Job& j, double* red, bool* finite = nullptr) {
  double s = 0.0;
  for (int i = threadIdx.x; i < j.sq_n; i += NT) s += j.sq_part[i];
  const double sq = block_sum_f64_waves<NT>(s, red);
  const double n = (double)j.nm;
  const double wmax = (double)j.wtab[j.nm / 2];
  const double bound = 2.0 * sq * wmax * wmax * wmax * n * n * n;
  if (finite) *finite = bound == bound && bound < 1e300;
  if (!(bound > 0.0) || !(bound < 1e300)) return 0;   // empty (or non-finite) field
  int ex;
  (void)frexp(bound, &ex);   // bound < 2^ex
  return 62 - ex;
}
__device__ __forceinline__ unsigned long long to_bin_fixed(float v, double scale) {
  return (unsigned long long)(long long)rint((double)v * scale);
}

// ------------------------------------------------------------------------------------
// x pass fused with deconvolution, |delta_k|^2, Legendre weights and per-bin FP64 sums.
// Each persistent block accumulates its own histogram and writes it to
// partials[block][nb][NACC] once at the end.
// ------------------------------------------------------------------------------------
// LDS of the binning x pass and the waves per SIMD it leaves room for; the register
// budget is held to that so registers never cost a resident block
template <int N> struct XBinLds {
  static constexpr size_t TILE = (size_t)N * (FFTCfg<N>::CX + 1) * 8;
  static constexpr int JLEN = isqrt_c(3 * (N / 2) * (N / 2)) + 2;   // isqrt-start table entries
  static constexpr size_t HIST =
      (size_t)(HB_LDS + 1) * 4 + 3 * HB_LDS * 8 + HB_LDS * 4 + (size_t)(N / 2 + 1) * 4 + (size_t)JLEN * 2 +
      (size_t)(N / 2) * 8;
  // LDS-limited waves per SIMD (the histogram in HBM: no limit is imposed)
  template <bool LH> static constexpr int wpe() {
    return LH ? cmax(1, cmin(4, (int)((160 * 1024) / (TILE + HIST)) * (FFTCfg<N>::NTX / 64) / 4)) : 1;
  }
};

template <int N, bool LH>
__device__ __forceinline__ void xpass_bin_body(const Job& j, const float2* __restrict__ twg, int bx, int nbx) {
  const float2* __restrict__ spec = j.spec;
  const float* __restrict__ wtab = j.wtab;
  const int* __restrict__ thr_g = j.thr;
  const int* __restrict__ jst_g = j.jst;
  const int nb = j.nb;
  using Cf = FFTCfg<N>;
  constexpr int C = Cf::CX, NT = Cf::NTX, PS = C + 1;
  constexpr int YT = N / C;
  constexpr int H = N / 2;
  // LDS: the tile, the S_0,2,4 histogram, the thresholds, the window factors of |k| <= N/2,
  // the isqrt-start table and the first half of the twiddles (<= 80 KB at N = 512 so two
  // blocks share a CU)
  constexpr int JLEN = XBinLds<N>::JLEN;
  __shared__ float2 buf[N * PS];
  __shared__ int thr_s[LH ? HB_LDS + 1 : 1];
  __shared__ unsigned long long hs[LH ? 3 * HB_LDS : 1];   // fixed-point S_0, S_2, S_4
  __shared__ double red_e[NT / 64 + 1];
  __shared__ unsigned hc[LH ? HB_LDS : 1];   // stored modes binned per bin
  __shared__ float wt_s[LH ? H + 1 : 1];
  __shared__ short jst_s[LH ? JLEN : 1];   // bins < HB_LDS
  __shared__ float2 twh[N / 2];
  const TwHalf<N> tw{twh};
  const int tid = threadIdx.x;
  for (int i = tid; i < N / 2; i += NT) twh[i] = twg[i];
  double* gpart = j.partials + (size_t)bx * nb * NACC;
  if constexpr (LH) {
    for (int i = tid; i <= nb; i += NT) thr_s[i] = thr_g[i];
    for (int i = tid; i < 3 * HB_LDS; i += NT) hs[i] = 0ull;
    for (int i = tid; i < HB_LDS; i += NT) hc[i] = 0u;
    for (int i = tid; i <= H; i += NT) wt_s[i] = wtab[i];
    for (int i = tid; i < JLEN; i += NT) jst_s[i] = (short)jst_g[i];
  } else {
    for (int i = tid; i < nb * NACC; i += NT) gpart[i] = 0.0;   // all-zero bits: integer 0 too
  }
  __syncthreads();
  const double fscale = ldexp(1.0, bin_fixed_exp<NT>(j, red_e));
  // threshold i (0 <= i <= nb) and the raw bin of m: #{i <= nb : thr[i] <= m} - 1, in
  // [-1, nb] (-1: below the first edge, nb: at or above the closed last edge)
  auto thr_at = [&](int i) -> int {
    if constexpr (LH) return thr_s[i];
    else return thr_g[i];
  };
  auto raw_bin = [&](int m) -> int {
    int jj = (int)sqrtf((float)m);
    if (jj * jj > m) --jj;
    else if ((jj + 1) * (jj + 1) <= m) ++jj;
    int bb;
    if constexpr (LH) bb = jst_s[jj];
    else bb = jst_g[jj];
    while (bb + 1 <= nb && thr_at(bb + 1) <= m) ++bb;
    return bb;
  };
  // window factor of array index i (symmetric: index N - i has the same |k|)
  auto wt = [&](int i) -> float {
    if constexpr (LH) return wt_s[i <= H ? i : N - i];
    else return wtab[i];
  };
  // Tasks.  Entries (q, ky) and (N - q, ky) share m and mu, and so do (q, -ky) and
  // (N - q, -ky).  With QUAD (C >= 16) a tile holds two column strips: S_a = ky in
  // [HW yt, HW yt + HW) and its mirror S_b = [N - HW yt - HW, N - HW yt) (HW = C / 2), so one
  // task bins the four entries of (q, +-ky).  Strip columns whose mirror lies in another
  // tile (ky = HW yt and ky = N - HW (yt + 1), the latter -|ky| with |ky| = HW (yt + 1)) are
  // single-column tasks, which also covers the self-mirrored ky = 0 and ky = N / 2.  There
  // are LPR = HW + 1 tasks per tile: task s < HW is strip column s (paired with S_b column
  // HW - s for s >= 1), task HW is S_b column 0.  Without QUAD (C < 16) a task is one column
  // and the pair (q, N - q).
  // Each lane walks a run of SL consecutive rows q of one task.  Along the run m = q^2 + r2
  // grows, so its bin only moves forward: one compare per row against the next threshold
  // tracks it, and the sums of a bin stay in registers until the bin changes or the run
  // ends, when they are added to the block's histogram.  The lanes of a wave take
  // consecutive runs of one task, so the bins they add to at the same step differ.
  constexpr bool QUAD = C >= 16;
  constexpr int HW = C / 2;
  constexpr int LPR = QUAD ? HW + 1 : C;          // tasks per tile
  constexpr int SEGS = NT / LPR;                  // runs per task
  static_assert(SEGS >= 1, "at least one run per task");
  constexpr int SL = (H + 1 + SEGS - 1) / SEGS;   // rows per run
  const int s = tid / SEGS < LPR ? tid / SEGS : LPR;   // s == LPR: idle lane
  const int qa = (tid % SEGS) * SL;                    // first row of the run
  const int qe = s < LPR ? min(qa + SL, H + 1) : qa;   // end of the run (exclusive)
  // global ky column of tile column c in column tile yt
  auto gcol = [](int c, int ytile) -> int {
    if constexpr (QUAD) return c < HW ? HW * ytile + c : N - HW * ytile - 2 * HW + c;
    else return ytile * C + c;
  };
  // N = 128: the next tile's columns are loaded into registers while this one is transformed
  // and binned (for larger N the registers would spill or cost a block; at N = 64 a block
  // has a single tile)
  constexpr bool XPF = LH && N == 128;
  // tile loads: every lane keeps one column (NT is a multiple of C) and rows r0 + k RSTEP,
  // so its loads share one base address and all of them are in flight before any use
  static_assert(NT % C == 0, "tile lanes keep one column");
  constexpr int RSTEP = NT / C;
  constexpr int XPER = (N + RSTEP - 1) / RSTEP;
  const int lc = tid % C, lr0 = tid / C;
  float2 xpre[XPF ? XPER : 1];
  auto load_tile = [&](int t, float2* v) {
    const float2* col = spec + (size_t)(t / YT) * N * N + gcol(lc, t % YT);
#pragma unroll
    for (int k = 0; k < XPER; ++k) {
      const int r = lr0 + k * RSTEP;
      if (N % RSTEP == 0 || r < N) v[k] = col[(size_t)r * N];
    }
  };
  auto store_tile = [&](const float2* v) {
#pragma unroll
    for (int k = 0; k < XPER; ++k) {
      const int r = lr0 + k * RSTEP;
      if (N % RSTEP == 0 || r < N) buf[r * PS + lc] = v[k];
    }
  };
  // First radix stage in registers: when RSTEP divides N / R0, a lane's rows lr0 + k RSTEP
  // are the inputs (rows j + r N / R0) of whole first-stage butterflies j = lr0 + m RSTEP, so
  // it transforms them and stores the stage's outputs (rows j R0 + r): the tile store, the
  // first stage's LDS read and one barrier fold into a single pass.
  constexpr int R0 = radix_at(N, 0);
  constexpr int B0 = N / R0;
  constexpr bool S0REG = N % RSTEP == 0 && B0 % RSTEP == 0;
  auto store_stage0 = [&](const float2* v) {
    constexpr int MB = S0REG ? B0 / RSTEP : 1;   // butterflies per lane
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
      float2 x[R0];
#pragma unroll
      for (int r = 0; r < R0; ++r) x[r] = v[mb + r * MB];   // row lr0 + (mb + r MB) RSTEP
      DFT<R0>::run(x);
      const int jb = lr0 + mb * RSTEP;
#pragma unroll
      for (int r = 0; r < R0; ++r) buf[(jb * R0 + r) * PS + lc] = x[r];
    }
  };
  auto store_first = [&](const float2* v) {
    if constexpr (S0REG) store_stage0(v);
    else store_tile(v);
  };
  // Work items enumerate only the tiles that hold a binned mode.  A tile whose columns all
  // have min ky^2 + kz^2 >= thr[nb] (the end of the last bin) holds only modes outside every
  // bin; those are neither loaded nor transformed, and the remaining tiles are dealt out to
  // the blocks evenly.  Plane kz keeps the tiles yt < lo(kz) and yt >= YT - hi(kz): the
  // smallest |ky| of a tile grows with yt up to the middle and falls beyond it (QUAD: strip
  // A starts at |ky| = HW yt and strip B's |ky| exceeds it, so hi = 0).  The table of the
  // planes' first work items and lo(kz) sits in the padding column c = C of buf, which no
  // tile store, transform stage or binning read touches.
  // Strips narrower than a 128-byte line (QUAD with HW < 16): the blocks of work items t and
  // t + 8 run on the same XCD at about the same time, so they take work items 2P and 2P + 1,
  // the two strips of a line while the planes before hold an even number of tiles.  (Keeping
  // every pair aligned, by rounding lo up to even or by listing the odd planes' last tiles
  // after all pairs, measured 7 us slower at 512^3: 165 against 157.)
  constexpr bool PAIRS = QUAD && HW * 8 < 128 && YT % 2 == 0;
  constexpr int KZN = H + 1;                      // planes kz = 0..H
  constexpr int KPT = (KZN + NT - 1) / NT;        // planes per lane
  constexpr int NW = NT / 64;
  static_assert(KZN + 1 + NW <= N, "plane table fits the padding column");
  int2* const ktab = reinterpret_cast<int2*>(buf);
  auto kslot = [](int i) -> int { return i * PS + C; };
  {
    const int thr_end = thr_at(nb);
    // smallest |ky| over the columns of tile yt
    auto kmin_of = [](int ytt) -> int {
      if constexpr (QUAD) {
        return HW * ytt;
      } else {
        const int y0 = ytt * C, y1 = y0 + C - 1;
        return min(y0 <= H ? y0 : N, y1 > H ? N - y1 : N);
      }
    };
    int lov[KPT], cntv[KPT], tsum = 0;
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int kz = tid * KPT + k;
      int lo = 0, hi = 0;
      if (kz < KZN) {
        const int rem = thr_end - kz * kz;   // needed: kmin^2 < rem
        while (lo < YT && kmin_of(lo) * kmin_of(lo) < rem) ++lo;
        if constexpr (!QUAD) {
          while (hi < YT - lo && kmin_of(YT - 1 - hi) * kmin_of(YT - 1 - hi) < rem) ++hi;
        }

      }
      lov[k] = lo;
      cntv[k] = lo + hi;
      tsum += lo + hi;
    }
    const int lane = tid & 63, wid = tid >> 6;
    int incl = tsum;
#pragma unroll
    for (int o = 1; o < 64; o *= 2) {
      const int v = __shfl_up(incl, o, 64);
      if (lane >= o) incl += v;
    }
    if (lane == 63) ktab[kslot(KZN + 1 + wid)].x = incl;
    __syncthreads();
    int run = incl - tsum, total = 0;
    for (int w = 0; w < NW; ++w) {
      const int wt_w = ktab[kslot(KZN + 1 + w)].x;
      run += w < wid ? wt_w : 0;
      total += wt_w;
    }
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int kz = tid * KPT + k;
      if (kz < KZN) ktab[kslot(kz)] = make_int2(run, lov[k]);
      run += cntv[k];
    }
    if (tid == 0) ktab[kslot(KZN)] = make_int2(total, 0);
    __syncthreads();
  }
  const int NC = ktab[kslot(KZN)].x;   // work items
  // tile of work item t (block-uniform): the plane is the last one whose first item is <= c
  auto tile_of = [&](int t) -> int {
    int c = t;
    if constexpr (PAIRS) {
      if (t < (NC / 16) * 16) c = 2 * ((t / 16) * 8 + (t % 8)) + ((t % 16) / 8);
    }
    int lo_i = 0, hi_i = KZN;   // first(lo_i) <= c < first(hi_i)
    while (hi_i - lo_i > 1) {
      const int mid = (lo_i + hi_i) >> 1;
      if (ktab[kslot(mid)].x <= c) lo_i = mid;
      else hi_i = mid;
    }
    const int2 e = ktab[kslot(lo_i)];
    const int jj = c - e.x;
    return lo_i * YT + (jj < e.y ? jj : YT - 1 - (jj - e.y));
  };
  int item = bx;
  int tile = item < NC ? tile_of(item) : 0;
  if constexpr (XPF) {
    if (item < NC) load_tile(tile, xpre);
  }
  while (item < NC) {
    const int next = item + nbx;
    const int tile_next = next < NC ? tile_of(next) : 0;
    const int kz = tile / YT, yt = tile % YT;
    const int kz2 = kz * kz;
    // this lane's task: |ky| (tile columns below, after the transform)
    int kyt;
    if constexpr (QUAD) {
      kyt = s < HW ? HW * yt + s : HW * (yt + 1);
    } else {
      const int y = yt * C + (s < LPR ? s : 0);
      kyt = y <= H ? y : N - y;
    }
    const int r2 = kyt * kyt + kz2;
    // the run's first bin and the threshold that ends it do not depend on the transform:
    // their table reads are issued before the tile's load and FFT
    int b = qa < qe ? raw_bin(qa * qa + r2) : nb;
    int thr_next = b < nb ? thr_at(b + 1) : 0x7fffffff;
    const float wyz = wt(kyt) * wt(kz);
    if constexpr (XPF) {
      store_first(xpre);
    } else {
      float2 v[XPER];
      load_tile(tile, v);
      store_first(v);
    }
    __syncthreads();
    if constexpr (XPF) {
      if (next < NC) load_tile(tile_next, xpre);
    }
    if constexpr (S0REG) FFTStages<N, 1, R0, C, PS, 1, NT, true, 1, 0>::run(buf, tw);
    else lds_fft<N, C, PS, 1, NT, true, 1>(buf, tw);
    // tile column A of the task, and B for a quad
    const bool quad = QUAD && s >= 1 && s < HW;
    const int cA = s < LPR ? s : 0;
    const int cB = quad ? 2 * HW - s : cA;
    float a0 = 0.0f, a2 = 0.0f, a4 = 0.0f;   // sums of the current bin
    int cnt = 0, cur = -1;
    auto flush = [&]() {
      if (cnt == 0) return;
      const unsigned long long f0 = to_bin_fixed(a0, fscale);
      const unsigned long long f2 = to_bin_fixed(a2, fscale);
      const unsigned long long f4 = to_bin_fixed(a4, fscale);
      if constexpr (LH) {
        atomicAdd(&hs[cur], f0);
        atomicAdd(&hs[HB_LDS + cur], f2);
        atomicAdd(&hs[2 * HB_LDS + cur], f4);
        atomicAdd(&hc[cur], (unsigned)cnt);
      } else {
        unsigned long long* g = reinterpret_cast<unsigned long long*>(gpart) + (size_t)cur * NACC;
        atomicAdd(g + 0, f0);
        atomicAdd(g + 1, f2);
        atomicAdd(g + 2, f4);
        atomicAdd(g + 3, (unsigned long long)cnt);
      }
    };
    // Each row reads its entries unconditionally (rows clamped to H, columns B = A for
    // singles) and masks what it adds, so the loop has no divergent branch but the flush of
    // a finished bin; masked entries add exact zeros, so the sums keep their order.
    const float fquad = quad ? 1.0f : 0.0f;
#pragma unroll 1
    for (int i = 0; i < SL; ++i) {
      const int q = qa + i;
      const bool valid = q < qe;
      const int m = q * q + r2;
      while (valid && m >= thr_next) {   // the bin moves forward (rarely more than one step)
        ++b;
        thr_next = b < nb ? thr_at(b + 1) : 0x7fffffff;
      }
      const int slot = !valid ? cur : (b >= 0 && b < nb ? b : -1);
      if (slot != cur) {
        flush();
        cur = slot;
        a0 = a2 = a4 = 0.0f;
        cnt = 0;
      }
      // the k = 0 mode's power is added by the finalize from the FP64 sum of the mesh
      // (mu = 0 there, :37), and its count too
      const bool use = valid && slot >= 0 && m != 0;
      const int qc = min(q, H);
      const bool pair = qc != 0 && qc != H;
      const int qm = qc == 0 ? 0 : N - qc;
      const float2 av = buf[qc * PS + cA];
      const float2 bv = buf[qm * PS + cA];
      const float2 cv = buf[qc * PS + cB];
      const float2 dv = buf[qm * PS + cB];
      const float fpair = pair ? 1.0f : 0.0f;
      float p2 = (av.x * av.x + av.y * av.y) + fpair * (bv.x * bv.x + bv.y * bv.y);
      p2 += fquad * ((cv.x * cv.x + cv.y * cv.y) + fpair * (dv.x * dv.x + dv.y * dv.y));
      p2 = use ? p2 * (wt(qc) * wyz) : 0.0f;
      // Legendre weights from integer numerators (mu^2 = kz^2 / m):
      //   L2 = (3 kz^2 - m) / (2 m),  L4 = (35 kz^4 - 30 kz^2 m + 3 m^2) / (8 m^2)
      // so L2 is exactly 0 where 3 kz^2 = m; 1 / m to within 1 ulp (hardware reciprocal);
      // the L4 numerator in float has relative error ~1e-7 of its terms (m = 0, masked,
      // takes m = 1 so that no infinity meets the zero weight)
      const int mm = m > 0 ? m : 1;
      const float rm = __builtin_amdgcn_rcpf((float)mm);
      const float mf = (float)mm;
      const float l2 = 0.5f * (float)(3 * kz2 - mm) * rm;
      const float num4 = fmaf((float)kz2, (float)(35 * kz2 - 30 * mm), 3.0f * mf * mf);
      const float l4 = 0.125f * num4 * rm * rm;
      a0 += p2;
      a2 = fmaf(p2, l2, a2);
      a4 = fmaf(p2, l4, a4);
      cnt += use ? (quad ? 2 : 1) * (pair ? 2 : 1) : 0;
    }
    flush();
    __syncthreads();
    item = next;
    tile = tile_next;
  }
  if constexpr (LH) {   // the block's integer partials (as int64 words)
    for (int i = tid; i < nb; i += NT) {
      unsigned long long* g = reinterpret_cast<unsigned long long*>(gpart) + (size_t)i * NACC;
      g[0] = hs[i];
      g[1] = hs[HB_LDS + i];
      g[2] = hs[2 * HB_LDS + i];
      g[3] = (unsigned long long)hc[i];
    }
  }
}

template <int N, bool LH>
__global__ void __launch_bounds__(FFTCfg<N>::NTX) __attribute__((amdgpu_waves_per_eu(XBinLds<N>::template wpe<LH>())))
k_xpass_bin(Job j, const float2* __restrict__ twg) {
  xpass_bin_body<N, LH>(j, twg, blockIdx.x, gridDim.x);
}

template <int N, bool LH>
__global__ void __launch_bounds__(FFTCfg<N>::NTX) __attribute__((amdgpu_waves_per_eu(XBinLds<N>::template wpe<LH>())))
k_xpass_bin_jobs(const Job* __restrict__ jobs, const float2* __restrict__ twg) {
  const Job& j = jobs[blockIdx.y];
  if (!j.active) return;
  xpass_bin_body<N, LH>(j, twg, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------------------------------------
// reduce the per-block partials in a fixed order and normalise (src/correlations.py:49-54)
// ------------------------------------------------------------------------------------
constexpr int FIN_NT = 256;

__device__ __forceinline__ void bin_finalize_body(const Job& j, int G, int b) {
  __shared__ double red[FIN_NT];
  const double* __restrict__ partials = j.partials;
  const int nb = j.nb;
  const int* __restrict__ thr = j.thr;
  const double* __restrict__ dc_part = j.dc_part;
  const int Gz = j.Gz;
  const double* __restrict__ scale_stats = j.scale_stats;
  const double norm = j.norm, shot = j.shot;
  double* __restrict__ sums = j.sums;
  int64_t* __restrict__ counts = j.counts;
  float* __restrict__ pk = j.pk;
  float* __restrict__ nmodes = j.nmodes;
  double* __restrict__ kmean = (j.flags & PK_FLAG_KMEAN) ? j.kmean : nullptr;
  __shared__ long long redi[FIN_NT / 64 + 1];
  // integer partials of the x pass blocks (int64 words), summed exactly, then scaled by the
  // x pass's fixed-point exponent (recomputed from the same partials in the same order)
  bool finite = true;
  const double unit = ldexp(1.0, -bin_fixed_exp<FIN_NT>(j, red, &finite));
  long long part[NACC] = {0, 0, 0, 0};
  for (int g = threadIdx.x; g < G; g += FIN_NT) {
    const long long* p = reinterpret_cast<const long long*>(partials) + ((size_t)g * nb + b) * NACC;
#pragma unroll
    for (int a = 0; a < NACC; ++a) part[a] += p[a];
  }
  double acc[NACC];
#pragma unroll
  for (int a = 0; a < NACC; ++a) {
    const long long t = block_sum_i64_waves<FIN_NT>(part[a], redi);
    acc[a] = a < 3 ? (finite ? (double)t * unit : __builtin_nan("")) : (double)t;
  }
  double dcp = 0.0;
  for (int g = threadIdx.x; dc_part != nullptr && g < Gz; g += FIN_NT) dcp += dc_part[g];
  const double dc = dc_part != nullptr ? block_sum_f64_waves<FIN_NT>(dcp, red) : 0.0;
  if (threadIdx.x != 0) return;
  // spectrum of rho - c (fused assignment): delta_k = rho_k * n^3 / M for k != 0
  const double sc = scale_stats ? scale_stats[1] * scale_stats[1] : 1.0;
  // k = 0 lands in the last bin with thr[b] <= 0 (thresholds are non-decreasing), if kept;
  // the x pass skips it, so its count and (unless the contrast zeroed it) power enter here
  const bool dc_bin = thr[b] <= 0 && (b + 1 == nb || thr[b + 1] > 0) && thr[nb] > 0;
  if (dc_part != nullptr && dc_bin) {
    const double p0 = dc * dc;   // mu = 0: L2 = -1/2, L4 = 3/8
    acc[0] += p0;
    acc[1] += -0.5 * p0;
    acc[2] += 0.375 * p0;
  }
  acc[0] *= sc;
  acc[1] *= sc;
  acc[2] *= sc;
  // stored modes the x pass binned (integer-valued partials, summed exactly), as the
  // reference's unit-weight histogram on the same pass (src/correlations.py:48)
  const int64_t nmode = (int64_t)acc[3] + (dc_bin ? 1 : 0);
  const double cnt = (double)nmode;
  if (sums) {
    sums[b * 3 + 0] = acc[0];
    sums[b * 3 + 1] = acc[1];
    sums[b * 3 + 2] = acc[2];
  }
  if (counts) counts[b] = nmode;
  if (nmodes) nmodes[b] = (float)cnt;
  // 0/0 -> NaN for empty bins, as the reference (:50-52)
  pk[b * 3 + 0] = (float)(acc[0] / cnt * norm - shot);
  pk[b * 3 + 1] = (float)(acc[1] / cnt * 5.0 * norm);
  pk[b * 3 + 2] = (float)(acc[2] / cnt * 9.0 * norm);
  if (kmean) kmean[b] = j.ksum[b] / cnt;
}

__global__ void __launch_bounds__(FIN_NT) k_bin_finalize(Job j, int G) { bin_finalize_body(j, G, blockIdx.x); }

__global__ void __launch_bounds__(FIN_NT) k_bin_finalize_jobs(const Job* __restrict__ jobs, int G) {
  const Job& j = jobs[blockIdx.y];
  if (!j.active || (int)blockIdx.x >= j.nb) return;
  bin_finalize_body(j, G, blockIdx.x);
}

// ------------------------------------------------------------------------------------
// Adjoint of the binning (backward of P_l(k) with respect to the density contrast).
//
// Forward: P_lb = c_lb sum_{k in b} W_k L_l(mu_k) |delta_k|^2 over the stored half
// spectrum, c_lb = norm (2l+1) / N_b.  With g_lb = dLoss/dP_lb and
// G_k = W_k sum_l g_lb c_lb L_l(mu_k), dLoss/d delta(x) = Re sum_{half} A_k e^{+i k.x} with
// A_k = 2 G_k delta_k (delta_k = S spec_k; S = stats[1] when the spectrum is of rho - c).
// Three passes mirror the forward ones: x (forward FFT, times 2 G S, inverse FFT),
// y (inverse), z (half-spectrum to real rows).
// ------------------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(FFTCfg<N>::NTX)
k_xpass_adjoint(Job j, const float2* __restrict__ twg, const float* __restrict__ coef, float2* __restrict__ bspec) {
  using Cf = FFTCfg<N>;
  constexpr int C = Cf::CX, NT = Cf::NTX, PS = C + 1;
  constexpr int YT = N / C;
  constexpr int NTILES = (N / 2 + 1) * YT;
  constexpr int H = N / 2;
  __shared__ float2 buf[N * PS];
  const float2* __restrict__ spec = j.spec;
  const float* __restrict__ wt = j.wtab;
  const int* __restrict__ thr = j.thr;
  const int* __restrict__ jst = j.jst;
  const int nb = j.nb;
  const float sf = j.scale_stats ? (float)j.scale_stats[1] : 1.0f;
  const int tid = threadIdx.x;
  for (int tile = blockIdx.x; tile < NTILES; tile += gridDim.x) {
    const int kz = tile / YT, yt = tile % YT;
    stage_tile<N, C, PS, NT>(buf, spec + (size_t)kz * N * N + (size_t)yt * C, N);
    __syncthreads();
    lds_fft<N, C, PS, 1, NT, true, 1>(buf, twg);
    const float wz = wt[kz];
    const int kz2 = kz * kz;
    for (int i = tid; i < N * C; i += NT) {
      const int c = i % C, r = i / C;
      const int y = yt * C + c;
      const int kx = r <= H ? r : r - N;
      const int ky = y <= H ? y : y - N;
      const int m = kx * kx + ky * ky + kz2;
      float g = 0.0f;
      if (m > 0) {   // the k = 0 mode is exactly 0 in the forward (contrast applied)
        const int slot = find_bin(m, thr, jst, nb);
        if (slot >= 0) {
          const float rm = 1.0f / (float)m;
          const float l2 = 0.5f * (float)(3 * kz2 - m) * rm;
          const double num4 = (double)kz2 * (35.0 * (double)kz2 - 30.0 * (double)m) + 3.0 * (double)m * (double)m;
          const float l4 = 0.125f * (float)num4 * rm * rm;
          g = wt[r] * wt[y] * wz * (coef[3 * slot] + coef[3 * slot + 1] * l2 + coef[3 * slot + 2] * l4);
        }
      }
      const float sc = 2.0f * g * sf;
      const float2 v = buf[r * PS + c];
      buf[r * PS + c] = make_float2(sc * v.x, -sc * v.y);   // conj(A): inverse via the forward FFT
    }
    __syncthreads();
    lds_fft<N, C, PS, 1, NT, true, 1>(buf, twg);
    float2* ob = bspec + (size_t)kz * N * N + (size_t)yt * C;
    for (int i = tid; i < N * C; i += NT) {
      const float2 v = buf[(i / C) * PS + (i % C)];
      ob[(size_t)(i / C) * N + (i % C)] = make_float2(v.x, -v.y);
    }
    __syncthreads();
  }
}

// kz_lim: planes above it are not transformed (they are zero for the caller)
template <int N>
__global__ void __launch_bounds__(FFTCfg<N>::NTY) k_ypass_inverse(float2* __restrict__ spec,
                                                                  const float2* __restrict__ twg, int kz_lim) {
  ypass_body<N, true>(spec, twg, blockIdx.x, gridDim.x, (kz_lim + 1) * N / FFTCfg<N>::CY);
}

// Half-spectrum rows B[kz][x][y] (kz = 0..N/2, transformed in x and y) -> real rows
// u(z) = Re B_0 + Re B_{N/2} (-1)^z + sum_{0<k<N/2} Re(B_k e^{2 pi i k z / N}), i.e. the C2R
// transform of Y_0 = Re B_0, Y_{N/2} = Re B_{N/2}, Y_k = B_k / 2: an N/2-point complex
// inverse FFT of Z_k = (Y_k + conj Y_{N/2-k}) + i e^{2 pi i k / N} (Y_k - conj Y_{N/2-k}),
// whose outputs are u(2j) + i u(2j+1).
// planes kz > kz_lim are taken as zero (not read)
template <int N>
__global__ void __launch_bounds__(FFTCfg<N>::NTZ)
k_zpass_c2r(const float2* __restrict__ bspec, const float2* __restrict__ twg, float* __restrict__ out, int kz_lim) {
  using Cf = FFTCfg<N>;
  constexpr int NC = Cf::NC, C = Cf::CZ, NT = Cf::NTZ, CS = NC + 1;
  constexpr int YT = N / C;
  constexpr int NTILES = N * YT;
  __shared__ float2 buf[C * CS];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += NT) tw[i] = twg[i];
  for (int tile = blockIdx.x; tile < NTILES; tile += gridDim.x) {
    const int x = tile / YT, yt = tile % YT;
    {
      constexpr int TOT = (NC + 1) * C, PER = (TOT + NT - 1) / NT;
      float2 v[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int i = tid + k * NT;
        if (TOT % NT == 0 || i < TOT)
          v[k] = (i / C) <= kz_lim ? bspec[((size_t)(i / C) * N + x) * N + (size_t)yt * C + (i % C)]
                                   : make_float2(0.0f, 0.0f);
      }
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int i = tid + k * NT;
        if (TOT % NT == 0 || i < TOT) buf[(i % C) * CS + i / C] = v[k];
      }
    }
    __syncthreads();
    // pairs (k, NC - k), k = 0..NC/2: each thread reads and rewrites only its own pair
    // (row[NC] is read by the k = 0 pair and never written), so no barrier is needed
    for (int i = tid; i < (NC / 2 + 1) * C; i += NT) {
      const int c = i % C, k = i / C;
      const int kk = NC - k;
      float2* row = buf + c * CS;
      auto Y = [&](int q) {
        const float2 b = row[q];
        if (q == 0 || q == NC) return make_float2(b.x, 0.0f);
        return make_float2(0.5f * b.x, 0.5f * b.y);
      };
      const float2 yk = Y(k), ykk = Y(kk);
      auto zval = [&](const float2 a, const float2 bconj_src, int q) {
        // (a + conj b) + i e^{2 pi i q / N} (a - conj b), e^{2 pi i q / N} = conj tw[q]
        const float2 s = make_float2(a.x + bconj_src.x, a.y - bconj_src.y);
        const float2 d = make_float2(a.x - bconj_src.x, a.y + bconj_src.y);
        const float2 e = make_float2(tw[q].x, -tw[q].y);
        const float2 ed = cmul(e, d);
        return make_float2(s.x - ed.y, s.y + ed.x);
      };
      const float2 zk = zval(yk, ykk, k);
      const float2 zkk = zval(ykk, yk, kk);
      row[k] = make_float2(zk.x, -zk.y);                 // conj: inverse via the forward FFT
      if (kk != k && kk < NC) row[kk] = make_float2(zkk.x, -zkk.y);
    }
    __syncthreads();
    lds_fft<NC, C, 1, CS, NT, true, 2>(buf, tw);
    float2* orow = reinterpret_cast<float2*>(out + ((size_t)x * N + (size_t)yt * C) * N);
    for (int i = tid; i < C * NC; i += NT) {
      const int c = i / NC, jj = i % NC;
      const float2 v = buf[c * CS + jj];
      orow[(size_t)c * NC + jj] = make_float2(v.x, -v.y);   // (u(2j), u(2j+1))
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------
// inverse x pass of numpy.fft.irfftn: columns along kx at fixed (kz, ky), inverse DFT
// via the forward FFT of the conjugate.  The input is a stored half spectrum
// (SRC = C2R_SPECTRUM, spec_in) or modes generated in the tile (C2R_DELTA: the Gaussian
// field, C2R_PSI_Z: its Zel'dovich z displacement), so a generated spectrum is never
// written or read back.  irfftn's 1/n^3 and the doubling of the interior kz planes (the
// half-spectrum -> real rows step below sums Re B_k once per interior k) are folded in
// here; the y pass (inverse) and the z pass (C2R) follow.
// ------------------------------------------------------------------------------------
template <int N, int SRC>
__global__ void __launch_bounds__(FFTCfg<N>::NTX)
k_xpass_inverse(const float2* __restrict__ spec_in, float2* __restrict__ out, const float2* __restrict__ twg,
                FieldGen g, ShellMask sm, int kz_lim) {
  using Cf = FFTCfg<N>;
  constexpr int C = Cf::CX, NT = Cf::NTX, PS = C + 1;
  constexpr int YT = N / C;
  constexpr int NTILES = (N / 2 + 1) * YT;
  constexpr int H = N / 2;
  __shared__ float2 buf[N * PS];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += NT) tw[i] = twg[i];
  const float s_edge = (float)(1.0 / ((double)N * (double)N * (double)N));
  const float s_int = 2.0f * s_edge;
  const int ntiles = min(NTILES, (kz_lim + 1) * YT);   // planes above kz_lim: zero, not written
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int kz = tile / YT, yt = tile % YT;
    const float sc = (kz == 0 || kz == H) ? s_edge : s_int;
    if constexpr (SRC == C2R_SPECTRUM) {
      const float2* src = spec_in + (size_t)kz * N * N + (size_t)yt * C;
      constexpr int TOT = N * C, PER = (TOT + NT - 1) / NT;
      float2 v[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int i = tid + k * NT;
        if (TOT % NT == 0 || i < TOT) v[k] = src[(size_t)(i / C) * N + (i % C)];
      }
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int i = tid + k * NT;
        if (TOT % NT == 0 || i < TOT) buf[(i / C) * PS + (i % C)] = make_float2(sc * v[k].x, -sc * v[k].y);
      }
    } else if constexpr (SRC == C2R_SHELL || SRC == C2R_SHELL_ONE) {
      // the k-shell lo <= m < hi of the stored spectrum times the window (src/correlations.py:
      // 373-382), or the shell's indicator
      const float2* src = spec_in + (size_t)kz * N * N + (size_t)yt * C;
      const float wz = sm.wt[kz];
      for (int i = tid; i < N * C; i += NT) {
        const int r = i / C, c = i % C;
        const int y = yt * C + c;
        const int sx = r <= H ? r : r - N, sy = y <= H ? y : y - N;
        const int m = sx * sx + sy * sy + kz * kz;
        float2 v = make_float2(0.0f, 0.0f);
        if (m >= sm.lo && m < sm.hi) {
          if constexpr (SRC == C2R_SHELL) {
            const float2 t = src[(size_t)r * N + c];
            const float w = sm.wt[r] * sm.wt[y] * wz * sc;
            v = make_float2(w * t.x, -w * t.y);
          } else {
            v = make_float2(sc, 0.0f);
          }
        }
        buf[r * PS + c] = v;
      }
    } else {
      for (int i = tid; i < N * C; i += NT) {
        const int r = i / C, c = i % C;
        const float2 v = SRC == C2R_DELTA ? gauss_mode(g, r, yt * C + c, kz) : psi_z_mode(g, r, yt * C + c, kz);
        buf[r * PS + c] = make_float2(sc * v.x, -sc * v.y);
      }
    }
    __syncthreads();
    lds_fft<N, C, PS, 1, NT, true, 1>(buf, tw);
    float2* ob = out + (size_t)kz * N * N + (size_t)yt * C;
    for (int i = tid; i < N * C; i += NT) {
      const float2 v = buf[(i / C) * PS + (i % C)];
      ob[(size_t)(i / C) * N + (i % C)] = make_float2(v.x, -v.y);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------
// bispectrum (src/correlations.py:334-458): the first two shells' fields (inverse
// transforms of the shell's delta_k W and of its indicator) are stored and multiplied
// (delta1 delta2, I1 I2); the later shells' sums are taken in Fourier space against the
// spectra of those products (k_bispec_kspace), so they need no inverse transform.
// ------------------------------------------------------------------------------------
// products of the first two shells' fields in place (p <- p q) and sums of their squares
constexpr int PAIR_NT = 256;
__global__ void __launch_bounds__(PAIR_NT) k_bispec_pair(float* __restrict__ p, const float* __restrict__ q,
                                                         int64_t total, double* __restrict__ part_p,
                                                         double* __restrict__ part_q) {
  __shared__ double red[PAIR_NT];
  double sp = 0.0, sq = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * PAIR_NT + threadIdx.x; i < total; i += (int64_t)gridDim.x * PAIR_NT) {
    const float a = p[i], b = q[i];
    sp += (double)a * (double)a;
    sq += (double)b * (double)b;
    p[i] = a * b;
  }
  const double ta = block_sum_f64(sp, red, PAIR_NT);
  const double tb = block_sum_f64(sq, red, PAIR_NT);
  if (threadIdx.x == 0) {
    part_p[blockIdx.x] = ta;
    part_q[blockIdx.x] = tb;
  }
}

// P_i = sum delta_i^2 / sum I_i^2 (L/n^2)^3, B = sum delta1 delta2 delta3 / sum I1 I2 I3
// (L^2/n^3)^3, Q = B / (P1 P2 + P1 P3 + P2 P3) (src/correlations.py:392-452)
// Shell sums of the later shells in Fourier space (Parseval, half spectrum with weight 1 on
// the kz = 0, n/2 planes and 2 elsewhere): with D = delta_k W on shell i, P = rfftn(delta1
// delta2), Q = rfftn(I1 I2),
//   sum_x delta_i^2       = n^-3 sum |D|^2,        sum_x I_i^2     = n^-3 sum 1,
//   sum_x delta1 delta2 delta_i = n^-3 sum Re(D conj P),  sum_x I1 I2 I_i = n^-3 sum Re Q,
// so no shell needs its own inverse transform.  Shells s0 .. s0 + ns - 1 (ns <= BKK_CHUNK)
// per launch, register accumulators, fixed-order block sums into the partial slots.
constexpr int BKK_NT = 256;
constexpr int BKK_CHUNK = 8;
__global__ void __launch_bounds__(BKK_NT)
k_bispec_kspace(const float2* __restrict__ dspec, const float2* __restrict__ pspec, const float2* __restrict__ qspec,
                const float* __restrict__ wt, ShellChunk sh, int s0, int ns, int n, int kz_lim, BispecArgs a) {
  static_assert(BKK_CHUNK == 8, "ShellChunk holds 8 shells");
  __shared__ double red[BKK_NT / 64 + 1];
  int slo[BKK_CHUNK], shi[BKK_CHUNK];
  int mlo = 0x7fffffff, mhi = 0;
#pragma unroll
  for (int j = 0; j < BKK_CHUNK; ++j) {
    slo[j] = j < ns ? sh.lo[j] : 1;
    shi[j] = j < ns ? sh.hi[j] : 0;
    if (j < ns) {
      mlo = min(mlo, slo[j]);
      mhi = max(mhi, shi[j]);
    }
  }
  double acc[BKK_CHUNK][4];
#pragma unroll
  for (int j = 0; j < BKK_CHUNK; ++j)
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[j][t] = 0.0;
  const int h = n / 2;
  const int64_t nn = (int64_t)n * n;
  const int64_t total = (int64_t)(kz_lim + 1) * nn;
  for (int64_t i = (int64_t)blockIdx.x * BKK_NT + threadIdx.x; i < total; i += (int64_t)gridDim.x * BKK_NT) {
    const int kz = (int)(i / nn);
    const int r = (int)(i - (int64_t)kz * nn);
    const int kx = r / n, ky = r - kx * n;
    const int sx = kx <= h ? kx : kx - n, sy = ky <= h ? ky : ky - n;
    const int m = sx * sx + sy * sy + kz * kz;
    if (m < mlo || m >= mhi) continue;
    const float2 d = dspec[i];
    const float2 p = pspec[i];
    const float w = wt[kx] * wt[ky] * wt[kz];
    const double wh = (kz == 0 || kz == h) ? 1.0 : 2.0;
    const double dx = (double)(d.x * w), dy = (double)(d.y * w);
    const double d2 = wh * (dx * dx + dy * dy);
    const double tr = wh * (dx * (double)p.x + dy * (double)p.y);
    const double qq = qspec ? wh * (double)qspec[i].x : 0.0;
#pragma unroll
    for (int j = 0; j < BKK_CHUNK; ++j)
      if (m >= slo[j] && m < shi[j]) {
        acc[j][0] += d2;
        acc[j][1] += wh;
        acc[j][2] += tr;
        acc[j][3] += qq;
      }
  }
  const double inv = 1.0 / ((double)n * (double)n * (double)n);
#pragma unroll
  for (int j = 0; j < BKK_CHUNK; ++j)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const double v = block_sum_f64_waves<BKK_NT>(acc[j][t], red);
      __syncthreads();
      if (j < ns && threadIdx.x == 0)
        a.part[((size_t)(s0 + j) * BK_SLOTS + t) * BK_PART_STRIDE + blockIdx.x] = v * inv;
    }
}

constexpr int BKF_NT = 256;
__global__ void __launch_bounds__(BKF_NT) k_bispec_finalize(BispecArgs a, int n, int g_pair, int g_red) {
  __shared__ double red[BKF_NT / 64 + 1];
  const int sh = blockIdx.x;   // one block per shell (uniform exit)
  double v[BK_SLOTS];
  for (int t = 0; t < BK_SLOTS; ++t) {
    const int cnt = (sh < 2) ? (t < 2 ? g_pair : 0) : g_red;
    const double* pp = a.part + ((size_t)sh * BK_SLOTS + t) * BK_PART_STRIDE;
    double s = 0.0;   // fixed order: strided per thread, then the block's fixed-order sum
    for (int i = threadIdx.x; i < cnt; i += BKF_NT) s += pp[i];
    v[t] = block_sum_f64_waves<BKF_NT>(s, red);
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (a.i_sums) {   // indicator sums depend only on n and the shells: taken as given
    v[1] = a.i_sums[2 * sh];
    v[3] = sh >= 2 ? a.i_sums[2 * sh + 1] : 0.0;
  }
  for (int t = 0; t < BK_SLOTS; ++t) a.sums[sh * BK_SLOTS + t] = v[t];
  const double h = a.box / ((double)n * (double)n);
  const double norm_p = h * h * h;
  a.pk[sh] = (float)(v[0] / v[1] * norm_p);
}

__global__ void __launch_bounds__(64) k_bispec_bq(BispecArgs a, int n) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i + 2 >= a.nshell) return;
  const double h = a.box / ((double)n * (double)n);
  const double norm_p = h * h * h;
  const double g = a.box * a.box / ((double)n * (double)n * (double)n);
  const double norm_b = g * g * g;
  const double* S = a.sums;
  const double p1 = S[0] / S[1] * norm_p, p2 = S[BK_SLOTS] / S[BK_SLOTS + 1] * norm_p;
  const double* s3 = S + (size_t)(i + 2) * BK_SLOTS;
  const double p3 = s3[0] / s3[1] * norm_p;
  const double b = s3[2] / s3[3] * norm_b;
  a.bk[i] = (float)b;
  a.qk[i] = (float)(b / (p1 * p2 + p1 * p3 + p2 * p3));
}

// ------------------------------------------------------------------------------------
// configuration-space multipoles xi_l(s) (src/correlations.py:119-260):
//   x pass: forward DFT along x, |delta_k|^2 W_k times irfftn's 1/n^3 (2/n^3 on interior
//           kz planes), inverse DFT along kx; FP64 partials of their sum = xi(s = 0),
//           the largest value of the separation grid (|xi(s)| <= xi(0) for a spectrum >= 0)
//   y pass: inverse (k_ypass_inverse)
//   z pass: C2R rows binned by |s| on the fly (the n^3 separation grid is never stored):
//           each thread walks a run of z, sums in FP64 while the bin stays the same and
//           adds each run to 64-bit fixed-point LDS bins (integer sums do not depend on
//           order, so results are run-to-run identical); blocks add their bins to global
//           accumulators; S = 2^e with (n^3 2 xi(0)) S < 2^62 bounds every bin sum
// ------------------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(FFTCfg<N>::NTX)
k_xpass_power_inverse(float2* __restrict__ spec, const float2* __restrict__ twg, const float* __restrict__ wt,
                      double* __restrict__ xi0_part) {
  using Cf = FFTCfg<N>;
  constexpr int C = Cf::CX, NT = Cf::NTX, PS = C + 1;
  constexpr int YT = N / C;
  constexpr int NTILES = (N / 2 + 1) * YT;
  constexpr int H = N / 2;
  __shared__ float2 buf[N * PS];
  __shared__ float2 tw[N];
  __shared__ double red[NT];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += NT) tw[i] = twg[i];
  const double s_edge = 1.0 / ((double)N * (double)N * (double)N);
  double acc = 0.0;
  for (int tile = blockIdx.x; tile < NTILES; tile += gridDim.x) {
    const int kz = tile / YT, yt = tile % YT;
    const float sc = (float)((kz == 0 || kz == H) ? s_edge : 2.0 * s_edge);
    float2* base = spec + (size_t)kz * N * N + (size_t)yt * C;
    stage_tile<N, C, PS, NT>(buf, base, N);
    __syncthreads();
    lds_fft<N, C, PS, 1, NT, true, 1>(buf, tw);
    const float wz = wt[kz];
    for (int i = tid; i < N * C; i += NT) {
      const int c = i % C, r = i / C;
      const float2 v = buf[r * PS + c];
      const float pw = (v.x * v.x + v.y * v.y) * (wt[r] * wt[yt * C + c] * wz) * sc;
      acc += (double)pw;
      buf[r * PS + c] = make_float2(pw, 0.0f);   // real: its conjugate is itself
    }
    __syncthreads();
    lds_fft<N, C, PS, 1, NT, true, 1>(buf, tw);
    for (int i = tid; i < N * C; i += NT) {
      const float2 v = buf[(i / C) * PS + (i % C)];
      base[(size_t)(i / C) * N + (i % C)] = make_float2(v.x, -v.y);
    }
    __syncthreads();
  }
  const double t = block_sum_f64(acc, red, NT);
  if (tid == 0) xi0_part[blockIdx.x] = t;
}

// fixed-order xi(0) -> scale; clears the accumulators of this estimate
__global__ void __launch_bounds__(256) k_xi_scale(XiArgs a, int n) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < a.gx; i += 256) s += a.xi0_part[i];
  const double xi0 = block_sum_f64(s, red, 256);
  for (int i = threadIdx.x; i < 3 * a.nb; i += 256) a.fixed[i] = 0ull;
  for (int i = threadIdx.x; i < a.nb; i += 256) a.nanflag[i] = 0;
  if (threadIdx.x == 0) {
    const double bound = 2.0 * xi0 * (double)n * (double)n * (double)n;
    int e = 0;
    double S = 1.0;
    if (bound > 0.0 && bound < 1e300) {
      (void)frexp(bound, &e);          // bound < 2^e
      S = ldexp(1.0, 62 - e);
    }
    a.scale[0] = S;
    a.scale[1] = 1.0 / S;
  }
}

template <int N> struct XiBinCfg {
  static constexpr int NC = N / 2, C = FFTCfg<N>::CZ, NT = FFTCfg<N>::NTZ;
  // complex values (two z each) per run: the largest divisor of NC not above NC C / NT
  static constexpr int want = cmax(1, NC * C / NT);
  static constexpr int div_le(int v, int d) { return (v % d == 0) ? d : div_le(v, d - 1); }
  static constexpr int CH = div_le(NC, want);
  static constexpr int RUNS = C * (NC / CH);
};

template <int N>
__global__ void __launch_bounds__(FFTCfg<N>::NTZ)
k_zpass_c2r_bin(const float2* __restrict__ bspec, const float2* __restrict__ twg, XiArgs a) {
  using Cf = FFTCfg<N>;
  using BC = XiBinCfg<N>;
  constexpr int NC = Cf::NC, C = Cf::CZ, NT = Cf::NTZ, CS = NC + 1;
  constexpr int YT = N / C;
  constexpr int NTILES = N * YT;
  constexpr int H = N / 2;
  constexpr int CH = BC::CH;
  __shared__ float2 buf[C * CS];
  __shared__ float2 tw[N];
  extern __shared__ unsigned long long bins[];   // [3 nb]: dynamic LDS sized at launch
  __shared__ int thr_s[XI_MAX_BINS + 1];
  __shared__ int jst_s[isqrt_c(3 * H * H) + 2];
  const int tid = threadIdx.x;
  const int nb = a.nb;
  for (int i = tid; i < N; i += NT) tw[i] = twg[i];
  for (int i = tid; i < 3 * nb; i += NT) bins[i] = 0ull;
  for (int i = tid; i <= nb; i += NT) thr_s[i] = a.thr[i];
  for (int i = tid; i < a.jlen; i += NT) jst_s[i] = a.jst[i];
  const double S = a.scale[0];
  for (int tile = blockIdx.x; tile < NTILES; tile += gridDim.x) {
    const int x = tile / YT, yt = tile % YT;
    {
      constexpr int TOT = (NC + 1) * C, PER = (TOT + NT - 1) / NT;
      float2 v[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int i = tid + k * NT;
        if (TOT % NT == 0 || i < TOT) v[k] = bspec[((size_t)(i / C) * N + x) * N + (size_t)yt * C + (i % C)];
      }
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int i = tid + k * NT;
        if (TOT % NT == 0 || i < TOT) buf[(i % C) * CS + i / C] = v[k];
      }
    }
    __syncthreads();
    for (int i = tid; i < (NC / 2 + 1) * C; i += NT) {
      const int c = i % C, k = i / C;
      const int kk = NC - k;
      float2* row = buf + c * CS;
      auto Y = [&](int q) {
        const float2 b = row[q];
        if (q == 0 || q == NC) return make_float2(b.x, 0.0f);
        return make_float2(0.5f * b.x, 0.5f * b.y);
      };
      const float2 yk = Y(k), ykk = Y(kk);
      auto zval = [&](const float2 p, const float2 q2, int q) {
        const float2 sm = make_float2(p.x + q2.x, p.y - q2.y);
        const float2 d = make_float2(p.x - q2.x, p.y + q2.y);
        const float2 e = make_float2(tw[q].x, -tw[q].y);
        const float2 ed = cmul(e, d);
        return make_float2(sm.x - ed.y, sm.y + ed.x);
      };
      const float2 zk = zval(yk, ykk, k);
      const float2 zkk = zval(ykk, yk, kk);
      row[k] = make_float2(zk.x, -zk.y);
      if (kk != k && kk < NC) row[kk] = make_float2(zkk.x, -zkk.y);
    }
    __syncthreads();
    lds_fft<NC, C, 1, CS, NT, true, 2>(buf, tw);
    const int sx = x <= H ? x : x - N;
    for (int w = tid; w < BC::RUNS; w += NT) {
      const int c = w / (NC / CH), j0 = (w % (NC / CH)) * CH;
      const int y = yt * C + c;
      const int sy = y <= H ? y : y - N;
      const int mxy = sx * sx + sy * sy;
      // the run's m range (|s_z| is unimodal in z: largest at z = n/2): runs whose every
      // m is outside [thr[0], thr[nb]) are skipped without touching their values
      {
        const int zf = 2 * j0, zl = 2 * (j0 + CH) - 1;
        const int af = zf <= H ? zf : N - zf, al = zl <= H ? zl : N - zl;
        const int amin = j0 == 0 ? 0 : min(af, al);
        const int amax = (zf <= H && zl >= H) ? H : max(af, al);
        if (mxy + amin * amin >= thr_s[nb] || mxy + amax * amax < thr_s[0]) continue;
      }
      int cur = -1;
      double a0 = 0.0, a2 = 0.0, a4 = 0.0;
      auto flush = [&]() {
        if (cur >= 0) {
          atomicAdd(&bins[3 * cur], (unsigned long long)__double2ll_rn(a0 * S));
          atomicAdd(&bins[3 * cur + 1], (unsigned long long)__double2ll_rn(a2 * S));
          atomicAdd(&bins[3 * cur + 2], (unsigned long long)__double2ll_rn(a4 * S));
        }
        a0 = a2 = a4 = 0.0;
      };
      for (int q = 0; q < CH; ++q) {
        const float2 v = buf[c * CS + j0 + q];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int z = 2 * (j0 + q) + e;
          const float val = e ? -v.y : v.x;      // (u(2j), u(2j+1)) = conj of the FFT output
          const int sz = z <= H ? z : z - N;
          const int z2 = sz * sz;
          const int m = mxy + z2;
          if (!(cur >= 0 && thr_s[cur] <= m && m < thr_s[cur + 1])) {
            flush();
            cur = find_bin(m, thr_s, jst_s, nb);
          }
          if (cur < 0) continue;
          const double dv = (double)val;
          a0 += dv;
          if (m == 0) {
            a.nanflag[cur] = 1;     // the s = 0 cell: mu = 0 / 0 (src/correlations.py:163)
          } else {
            const float rm = __builtin_amdgcn_rcpf((float)m);
            const float l2 = 0.5f * (float)(3 * z2 - m) * rm;
            const double num4 = (double)z2 * (35.0 * (double)z2 - 30.0 * (double)m) + 3.0 * (double)m * (double)m;
            const float l4 = 0.125f * (float)num4 * rm * rm;
            a2 += dv * (double)l2;
            a4 += dv * (double)l4;
          }
        }
      }
      flush();
    }
    __syncthreads();
  }
  for (int i = tid; i < 3 * nb; i += NT)
    if (bins[i]) atomicAdd(&a.fixed[i], bins[i]);
}

// FP64 sums and xi_l = (2l+1) sum_l / (N n^3); empty bins per the variant
__global__ void __launch_bounds__(256) k_xi_finalize(XiArgs a, int n) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= a.nb) return;
  const double inv = a.scale[1];
  const double n3 = (double)n * (double)n * (double)n;
  const long long cnt = a.modes[b];
  const bool nan24 = a.nanflag[b] != 0;
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    const double sl = (double)(long long)a.fixed[3 * b + l] * inv;
    a.sums[3 * b + l] = sl;
    float v;
    if (cnt == 0) v = a.empty_nan ? __int_as_float(0x7fc00000) : 0.0f;
    else if (l > 0 && nan24) v = __int_as_float(0x7fc00000);
    else v = (float)(sl / (double)cnt * (double)(4 * l + 1) / n3);
    a.xi[3 * b + l] = v;
  }
  a.nmodes[b] = cnt == 0 ? (a.empty_nan ? 0.0f : __int_as_float(0x7f800000)) : (float)cnt;
}

// ------------------------------------------------------------------------------------
// host launchers
// ------------------------------------------------------------------------------------
template <typename K>
static int persistent_grid(K kernel, int nt, int ntiles) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  int cus = num_cus(dev);
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, nt, 0) != hipSuccess || occ < 1) occ = 1;
  long g = (long)cus * occ;
  if (g > ntiles) g = ntiles;
  return (int)(g < 1 ? 1 : g);
}

template <int N> struct FFTLaunch {
  using Cf = FFTCfg<N>;
  static constexpr int ZTILES = N * (N / Cf::CZ);
  static constexpr int YTILES = (N / 2 + 1) * N / Cf::CY;
  static constexpr int XTILES = (N / 2 + 1) * (N / Cf::CX);
  static int zgrid() {
    static int grid = 0;
    if (!grid) grid = persistent_grid(k_zpass<N>, Cf::NTZ, ZTILES);
    return grid;
  }
  static int ygrid() {
    static int grid = 0;
    if (!grid) grid = persistent_grid(k_ypass<N>, Cf::NTY, YTILES);
    return grid;
  }
  static int zpass(const float* mesh, float2* spec, const float2* tw, const double* stats, double* dc_part,
                   double* sq_part, hipStream_t s, int kz_lim = N / 2) {
    const int grid = zgrid();
    StageTimer t(ST_ZPASS, s);
    hipLaunchKernelGGL(k_zpass<N>, dim3(grid), dim3(Cf::NTZ), 0, s, mesh, spec, tw, stats, dc_part, sq_part,
                       kz_lim);
    return check_hip(hipGetLastError(), "z pass launch");
  }
  static int ypass(float2* spec, const float2* tw, hipStream_t s) {
    const int grid = ygrid();
    StageTimer t(ST_YPASS, s);
    hipLaunchKernelGGL(k_ypass<N>, dim3(grid), dim3(Cf::NTY), 0, s, spec, tw);
    return check_hip(hipGetLastError(), "y pass launch");
  }
  static int ypass_mass(const Job& j, const float2* tw, hipStream_t s) {
    static int grid = 0;
    if (!grid) grid = persistent_grid(k_ypass_mass<N>, Cf::NTY, YTILES);
    StageTimer t(ST_YPASS, s);
    hipLaunchKernelGGL(k_ypass_mass<N>, dim3(grid), dim3(Cf::NTY), 0, s, j, tw);
    return check_hip(hipGetLastError(), "y pass launch");
  }
  static int ypass_jobs(const Job* jobs, int njobs, int gy, const float2* tw, hipStream_t s) {
    if (gy < 1) gy = 1;
    if (gy > YTILES) gy = YTILES;
    StageTimer t(ST_YPASS, s);
    hipLaunchKernelGGL(k_ypass_jobs<N>, dim3(gy, njobs), dim3(Cf::NTY), 0, s, jobs, tw);
    return check_hip(hipGetLastError(), "batched y pass launch");
  }
  static int xpass(float2* spec, const float2* tw, hipStream_t s, int kz_lim = N / 2) {
    static int grid = 0;
    if (!grid) grid = persistent_grid(k_xpass<N>, Cf::NTX, XTILES);
    StageTimer t(ST_XPASS_PLAIN, s);
    hipLaunchKernelGGL(k_xpass<N>, dim3(grid), dim3(Cf::NTX), 0, s, spec, tw, kz_lim);
    return check_hip(hipGetLastError(), "x pass launch");
  }
  // blocks per job of a batch of njobs: the resident capacity of the whole chip shared out,
  // at most one block per tile
  static int ypass_jobs_grid(int njobs) {
    static int cap = 0;
    if (!cap) cap = persistent_grid(k_ypass_jobs<N>, Cf::NTY, 1 << 30);
    return share(cap, njobs, YTILES);
  }
  static int bin_jobs_grid(bool lh, int njobs) {
    static int c_l = 0, c_g = 0;
    int& cap = lh ? c_l : c_g;
    if (!cap)
      cap = lh ? persistent_grid(k_xpass_bin_jobs<N, true>, Cf::NTX, 1 << 30)
               : persistent_grid(k_xpass_bin_jobs<N, false>, Cf::NTX, 1 << 30);
    return share(cap, njobs, bin_grid(lh));
  }
  static int share(int cap, int njobs, int limit) {
    int g = (cap + njobs - 1) / (njobs < 1 ? 1 : njobs);
    if (g > limit) g = limit;
    return g < 1 ? 1 : g;
  }
  static int bin_grid(bool lh) {
    static int g_l = 0, g_g = 0;
    if (lh) {
      if (!g_l) g_l = persistent_grid(k_xpass_bin<N, true>, Cf::NTX, XTILES);
      return g_l;
    }
    if (!g_g) g_g = persistent_grid(k_xpass_bin<N, false>, Cf::NTX, XTILES);
    return g_g;
  }
  static int xpass_bin(const Job& j, const float2* tw, int grid, hipStream_t s) {
    StageTimer t(ST_XPASS_BIN, s);
    if (j.nb <= HB_LDS)
      hipLaunchKernelGGL((k_xpass_bin<N, true>), dim3(grid), dim3(Cf::NTX), 0, s, j, tw);
    else
      hipLaunchKernelGGL((k_xpass_bin<N, false>), dim3(grid), dim3(Cf::NTX), 0, s, j, tw);
    return check_hip(hipGetLastError(), "x pass + bin launch");
  }
  static int adjoint(const Job& j, const float2* tw, const float* coef, float2* bspec, float* gmesh, hipStream_t s) {
    static int gx = 0, gy = 0, gz = 0;
    if (!gx) gx = persistent_grid(k_xpass_adjoint<N>, Cf::NTX, XTILES);
    if (!gy) gy = persistent_grid(k_ypass_inverse<N>, Cf::NTY, YTILES);
    if (!gz) gz = persistent_grid(k_zpass_c2r<N>, Cf::NTZ, ZTILES);
    StageTimer t(ST_ADJOINT_FFT, s);
    hipLaunchKernelGGL(k_xpass_adjoint<N>, dim3(gx), dim3(Cf::NTX), 0, s, j, tw, coef, bspec);
    hipLaunchKernelGGL(k_ypass_inverse<N>, dim3(gy), dim3(Cf::NTY), 0, s, bspec, tw, N / 2);
    hipLaunchKernelGGL(k_zpass_c2r<N>, dim3(gz), dim3(Cf::NTZ), 0, s, (const float2*)bspec, tw, gmesh, N / 2);
    return check_hip(hipGetLastError(), "adjoint FFT launch");
  }
  template <int SRC>
  static int xinv(const float2* in, float2* scratch, const float2* tw, const FieldGen& g, hipStream_t s,
                  ShellMask sm = ShellMask{0, 0, nullptr}, int kz_lim = N / 2) {
    static int gx = 0;
    if (!gx) gx = persistent_grid(k_xpass_inverse<N, SRC>, Cf::NTX, XTILES);
    hipLaunchKernelGGL((k_xpass_inverse<N, SRC>), dim3(gx), dim3(Cf::NTX), 0, s, in, scratch, tw, g, sm, kz_lim);
    return check_hip(hipGetLastError(), "inverse x pass launch");
  }
  static int kspace_grid() {
    int dev = 0;
    (void)hipGetDevice(&dev);
    return cmin(BK_PART_STRIDE, num_cus(dev) * 8);
  }
  static int pair_grid() {
    static int g = 0;
    if (!g) {
      int dev = 0;
      (void)hipGetDevice(&dev);
      g = cmin(BK_PART_STRIDE, num_cus(dev) * 8);
    }
    return g;
  }
  // shell inverse: x pass (masked source) + inverse y, into scratch
  // last plane a shell lo <= m < hi can touch: kz^2 <= m < hi
  static int shell_kz_lim(int hi) {
    if (hi <= 0) return -1;
    int k = 0;
    while ((k + 1) * (k + 1) <= hi - 1) ++k;
    return k < N / 2 ? k : N / 2;
  }
  static int shell_inverse(bool one, const float2* spec, float2* scratch, const float2* tw, ShellMask sm,
                           hipStream_t s) {
    static int gy = 0;
    if (!gy) gy = persistent_grid(k_ypass_inverse<N>, Cf::NTY, YTILES);
    FieldGen g{};
    g.n = N;
    const int kl = shell_kz_lim(sm.hi);
    int e = one ? xinv<C2R_SHELL_ONE>(spec, scratch, tw, g, s, sm, kl) : xinv<C2R_SHELL>(spec, scratch, tw, g, s, sm, kl);
    if (e) return e;
    if (kl >= 0) hipLaunchKernelGGL(k_ypass_inverse<N>, dim3(gy), dim3(Cf::NTY), 0, s, scratch, tw, kl);
    return check_hip(hipGetLastError(), "shell inverse launch");
  }
  // forward transform of the planes kz <= kz_lim only (higher planes are left unwritten)
  static int r2c_into(const float* mesh, float2* out, const float2* tw, hipStream_t s, int kz_lim = N / 2) {
    static int gy = 0;
    if (!gy) gy = persistent_grid(k_ypass_lim<N>, Cf::NTY, YTILES);
    int e = zpass(mesh, out, tw, nullptr, nullptr, nullptr, s, kz_lim);
    if (e) return e;
    if (kz_lim >= 0) hipLaunchKernelGGL(k_ypass_lim<N>, dim3(gy), dim3(Cf::NTY), 0, s, out, tw, kz_lim);
    e = check_hip(hipGetLastError(), "y pass launch");
    if (!e) e = xpass(out, tw, s, kz_lim);
    return e;
  }
  static int bispec(const float* mesh, float2* spec, float2* scratch, float2* qspec, float* const* meshes,
                    const float2* tw, const float* wt, const int* lo, const int* hi, const BispecArgs& a,
                    hipStream_t s) {
    static int gz = 0;
    if (!gz) gz = persistent_grid(k_zpass_c2r<N>, Cf::NTZ, ZTILES);
    // no shell reads a plane above the largest shell's limit
    int klmax = -1;
    for (int sh = 0; sh < a.nshell; ++sh) klmax = klmax > shell_kz_lim(hi[sh]) ? klmax : shell_kz_lim(hi[sh]);
    int e = r2c_into(mesh, spec, tw, s, klmax);
    if (e) return e;
    // first two shells: stored fields (the indicator fields only when their sums are unknown)
    const int nones = a.i_sums ? 1 : 2;
    for (int sh = 0; sh < 2; ++sh)
      for (int one = 0; one < nones; ++one) {
        e = shell_inverse(one != 0, spec, scratch, tw, ShellMask{lo[sh], hi[sh], wt}, s);
        if (e) return e;
        hipLaunchKernelGGL(k_zpass_c2r<N>, dim3(gz), dim3(Cf::NTZ), 0, s, (const float2*)scratch, tw,
                           meshes[one * 2 + sh], shell_kz_lim(hi[sh]));
      }
    const int gp = pair_grid();
    const int64_t total = (int64_t)N * N * N;
    auto slot = [&](int sh, int t) { return a.part + ((size_t)sh * BK_SLOTS + t) * BK_PART_STRIDE; };
    hipLaunchKernelGGL(k_bispec_pair, dim3(gp), dim3(PAIR_NT), 0, s, meshes[0], (const float*)meshes[1], total,
                       slot(0, 0), slot(1, 0));
    if (!a.i_sums)
      hipLaunchKernelGGL(k_bispec_pair, dim3(gp), dim3(PAIR_NT), 0, s, meshes[2], (const float*)meshes[3], total,
                         slot(0, 1), slot(1, 1));
    // later shells in Fourier space against the spectra of the product meshes
    int gr = 1;
    if (a.nshell > 2) {
      int kl = -1;
      for (int sh = 2; sh < a.nshell; ++sh) kl = kl > shell_kz_lim(hi[sh]) ? kl : shell_kz_lim(hi[sh]);
      e = r2c_into(meshes[0], scratch, tw, s, kl);
      if (!e && !a.i_sums) e = r2c_into(meshes[2], qspec, tw, s, kl);
      if (e) return e;
      gr = kspace_grid();
      for (int s0 = 2; s0 < a.nshell; s0 += BKK_CHUNK) {
        const int ns = a.nshell - s0 < BKK_CHUNK ? a.nshell - s0 : BKK_CHUNK;
        ShellChunk ch{};
        for (int j = 0; j < ns; ++j) {
          ch.lo[j] = lo[s0 + j];
          ch.hi[j] = hi[s0 + j];
        }
        hipLaunchKernelGGL(k_bispec_kspace, dim3(gr), dim3(BKK_NT), 0, s, (const float2*)spec, (const float2*)scratch,
                           a.i_sums ? nullptr : (const float2*)qspec, wt, ch, s0, ns, N, kl, a);
      }
    }
    hipLaunchKernelGGL(k_bispec_finalize, dim3(a.nshell), dim3(BKF_NT), 0, s, a, N, gp, gr);
    if (a.nshell > 2)
      hipLaunchKernelGGL(k_bispec_bq, dim3((a.nshell - 2 + 63) / 64), dim3(64), 0, s, a, N);
    return check_hip(hipGetLastError(), "bispectrum launch");
  }
  static int c2r(int src, const float2* in, float2* scratch, const FieldGen& g, const float2* tw, float* mesh,
                 hipStream_t s) {
    static int gy = 0, gz = 0;
    if (!gy) gy = persistent_grid(k_ypass_inverse<N>, Cf::NTY, YTILES);
    if (!gz) gz = persistent_grid(k_zpass_c2r<N>, Cf::NTZ, ZTILES);
    int e = src == C2R_SPECTRUM ? xinv<C2R_SPECTRUM>(in, scratch, tw, g, s)
            : src == C2R_DELTA  ? xinv<C2R_DELTA>(in, scratch, tw, g, s)
                                : xinv<C2R_PSI_Z>(in, scratch, tw, g, s);
    if (e) return e;
    hipLaunchKernelGGL(k_ypass_inverse<N>, dim3(gy), dim3(Cf::NTY), 0, s, scratch, tw, N / 2);
    hipLaunchKernelGGL(k_zpass_c2r<N>, dim3(gz), dim3(Cf::NTZ), 0, s, (const float2*)scratch, tw, mesh, N / 2);
    return check_hip(hipGetLastError(), "inverse FFT launch");
  }
  static int xi_gx() {
    static int gx = 0;
    if (!gx) gx = persistent_grid(k_xpass_power_inverse<N>, Cf::NTX, XTILES);
    return gx;
  }
  static int xi(const float* mesh, float2* spec, const float2* tw, const float* wt, const XiArgs& a,
                hipStream_t s) {
    static int gy = 0, gz = 0;
    if (!gy) gy = persistent_grid(k_ypass_inverse<N>, Cf::NTY, YTILES);
    if (!gz) gz = persistent_grid(k_zpass_c2r_bin<N>, Cf::NTZ, ZTILES);   // sized without dynamic LDS
    int e = zpass(mesh, spec, tw, nullptr, nullptr, nullptr, s);
    if (!e) e = ypass(spec, tw, s);
    if (e) return e;
    hipLaunchKernelGGL(k_xpass_power_inverse<N>, dim3(xi_gx()), dim3(Cf::NTX), 0, s, spec, tw, wt, a.xi0_part);
    hipLaunchKernelGGL(k_xi_scale, dim3(1), dim3(256), 0, s, a, N);
    hipLaunchKernelGGL(k_ypass_inverse<N>, dim3(gy), dim3(Cf::NTY), 0, s, spec, tw, N / 2);
    hipLaunchKernelGGL(k_zpass_c2r_bin<N>, dim3(gz), dim3(Cf::NTZ), (size_t)a.nb * 3 * sizeof(unsigned long long), s,
                       (const float2*)spec, tw, a);
    hipLaunchKernelGGL(k_xi_finalize, dim3((a.nb + 255) / 256), dim3(256), 0, s, a, N);
    return check_hip(hipGetLastError(), "xi launch");
  }
  static int xpass_bin_jobs(const Job* jobs, int njobs, int max_nb, int gx, const float2* tw, hipStream_t s) {
    StageTimer t(ST_XPASS_BIN, s);
    if (max_nb <= HB_LDS)
      hipLaunchKernelGGL((k_xpass_bin_jobs<N, true>), dim3(gx, njobs), dim3(Cf::NTX), 0, s, jobs, tw);
    else
      hipLaunchKernelGGL((k_xpass_bin_jobs<N, false>), dim3(gx, njobs), dim3(Cf::NTX), 0, s, jobs, tw);
    return check_hip(hipGetLastError(), "batched x pass + bin launch");
  }
};

#define PK_FOR_EACH_N(X) \
  X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)

bool fft_supported(int n) {
  switch (n) {
#define PK_CASE(V) case V: return true;
    PK_FOR_EACH_N(PK_CASE)
#undef PK_CASE
    default: return false;
  }
}

int zpass_grid(int n) {
  switch (n) {
#define PK_CASE(V) case V: return FFTLaunch<V>::zgrid();
    PK_FOR_EACH_N(PK_CASE)
#undef PK_CASE
  }
  return 0;
}

int ypass_grid(int n) {
  switch (n) {
#define PK_CASE(V) case V: return FFTLaunch<V>::ygrid();
    PK_FOR_EACH_N(PK_CASE)
#undef PK_CASE
  }
  return 0;
}

int fft_zpass(int n, const float* mesh, float2* spec, const float2* tw, const double* stats, double* dc_part,
              double* sq_part, hipStream_t s) {
  switch (n) {
#define PK_CASE(V) case V: return FFTLaunch<V>::zpass(mesh, spec, tw, stats, dc_part, sq_part, s);
    PK_FOR_EACH_N(PK_CASE)
#undef PK_CASE
  }
  return set_error(PK_E_INVALID, "unsupported n=%d", n);
}

int fft_ypass(int n, const Job& j, const float2* tw, hipStream_t s) {
  switch (n) {
#define PK_CASE(V) case V: return FFTLaunch<V>::ypass(j.spec, tw, s);
    PK_FOR_EACH_N(PK_CASE)
#undef PK_CASE
  }
  return set_error(PK_E_INVALID, "unsupported n=%d", n);
}

int fft_ypass_mass(int n, const Job& j, const float2* tw, hipStream_t s) {
  switch (n) {
#define PK_CASE(V) case V: return FFTLaunch<V>::ypass_mass(j, tw, s);
    PK_FOR_EACH_N(PK_CASE)
#undef PK_CASE
  }
  return set_error(PK_E_INVALID, "unsupported n=%d", n);
}

int fft_ypass_jobs(int n, const Job* jobs, int njobs, int gy, const float2* tw, hipStream_t s) {
  switch (n) {
#define PK_CASE(V) case V: return FFTLaunch<V>::ypass_jobs(jobs, njobs, gy, tw, s);
    PK_FOR_EACH_N(PK_CASE)
#undef PK_CASE
  }
  return set_error(PK_E_INVALID, "unsupported n=%d", n);
}

int fft_xpass(int n, float2* spec, const float2* tw, hipStream_t s) {
  switch (n) {
#define PK_CASE(V) case V: return FFTLaunch<V>::xpass(spec, tw, s);
    PK_FOR_EACH_N(PK_CASE)
#undef PK_CASE
  }
  return set_error(PK_E_INVALID, "unsupported n=%d", n);
}

int ypass_jobs_grid(int n, int njobs) {
  switch (n) {
#define PK_CASE(V) case V: return FFTLaunch<V>::ypass_jobs_grid(njobs);
    PK_FOR_EACH_N(PK_CASE)
#undef PK_CASE
  }
  return 1;
}

int bin_jobs_grid(int n, int nb, int njobs) {
  switch (n) {
#define PK_CASE(V) case V: return FFTLaunch<V>::bin_jobs_grid(nb <= HB_LDS, njobs);
    PK_FOR_EACH_N(PK_CASE)
#undef PK_CASE
  }
  return 1;
}

int bin_grid(int n, int nb) {
  switch (n) {
#define PK_CASE(V) case V: return FFTLaunch<V>::bin_grid(nb <= HB_LDS);
    PK_FOR_EACH_N(PK_CASE)
#undef PK_CASE
  }
  return 0;
}

int fft_xpass_bin(int n, const Job& j, const float2* tw, int grid, hipStream_t s) {
  switch (n) {
#define PK_CASE(V) case V: return FFTLaunch<V>::xpass_bin(j, tw, grid, s);
    PK_FOR_EACH_N(PK_CASE)
#undef PK_CASE
  }
  return set_error(PK_E_INVALID, "unsupported n=%d", n);
}

int fft_xpass_bin_jobs(int n, const Job* jobs, int njobs, int max_nb, int gx, const float2* tw, hipStream_t s) {
  switch (n) {
#define PK_CASE(V) case V: return FFTLaunch<V>::xpass_bin_jobs(jobs, njobs, max_nb, gx, tw, s);
    PK_FOR_EACH_N(PK_CASE)
#undef PK_CASE
  }
  return set_error(PK_E_INVALID, "unsupported n=%d", n);
}

int fft_adjoint(int n, const Job& j, const float2* tw, const float* coef, float2* bspec, float* gmesh, hipStream_t s) {
  switch (n) {
#define PK_CASE(V) case V: return FFTLaunch<V>::adjoint(j, tw, coef, bspec, gmesh, s);
    PK_FOR_EACH_N(PK_CASE)
#undef PK_CASE
  }
  return set_error(PK_E_INVALID, "unsupported n=%d", n);
}

int fft_c2r(int n, int src, const float2* spec_in, float2* scratch, const FieldGen& g, const float2* tw, float* mesh,
            hipStream_t s) {
  switch (n) {
#define PK_CASE(V) case V: return FFTLaunch<V>::c2r(src, spec_in, scratch, g, tw, mesh, s);
    PK_FOR_EACH_N(PK_CASE)
#undef PK_CASE
  }
  return set_error(PK_E_INVALID, "unsupported n=%d", n);
}

int xi_grid_x(int n) {
  switch (n) {
#define PK_CASE(V) case V: return FFTLaunch<V>::xi_gx();
    PK_FOR_EACH_N(PK_CASE)
#undef PK_CASE
  }
  return 0;
}

int fft_bispec(int n, const float* mesh, float2* spec, float2* scratch, float2* qspec, float* const* meshes,
               const float2* tw, const float* wt, const int* lo, const int* hi, const BispecArgs& a, hipStream_t s) {
  switch (n) {
#define PK_CASE(V) case V: return FFTLaunch<V>::bispec(mesh, spec, scratch, qspec, meshes, tw, wt, lo, hi, a, s);
    PK_FOR_EACH_N(PK_CASE)
#undef PK_CASE
  }
  return set_error(PK_E_INVALID, "unsupported n=%d", n);
}

int fft_xi(int n, const float* mesh, float2* spec, const float2* tw, const float* wt, const XiArgs& a,
           hipStream_t s) {
  switch (n) {
#define PK_CASE(V) case V: return FFTLaunch<V>::xi(mesh, spec, tw, wt, a, s);
    PK_FOR_EACH_N(PK_CASE)
#undef PK_CASE
  }
  return set_error(PK_E_INVALID, "unsupported n=%d", n);
}

int bin_finalize(const Job& j, int G, hipStream_t s) {
  StageTimer t(ST_FINALIZE, s);
  hipLaunchKernelGGL(k_bin_finalize, dim3(j.nb), dim3(FIN_NT), 0, s, j, G);
  return check_hip(hipGetLastError(), "bin finalize launch");
}

int bin_finalize_jobs(const Job* jobs, int njobs, int max_nb, int G, hipStream_t s) {
  StageTimer t(ST_FINALIZE, s);
  hipLaunchKernelGGL(k_bin_finalize_jobs, dim3(max_nb, njobs), dim3(FIN_NT), 0, s, jobs, G);
  return check_hip(hipGetLastError(), "batched bin finalize launch");
}

}  // namespace pk
