// Shared device helpers: complex arithmetic, small-radix DFTs and the LDS-resident
// mixed-radix Stockham stages used by every FFT pass.  gfx950 (wave64) only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pk_api.h"

namespace pk {

// ------------------------------------------------------------------------------------
// host-side error state (defined in pk_capi.hip)
// ------------------------------------------------------------------------------------
int set_error(int code, const char* fmt, ...);
int check_hip(hipError_t e, const char* what);

enum Stage {
  ST_SORT_COUNT = 0, ST_SCAN, ST_SORT_PLACE, ST_DEPOSIT, ST_ATOMIC, ST_MASS,
  ST_ZPASS, ST_YPASS, ST_XPASS_BIN, ST_FINALIZE, ST_XPASS_PLAIN, ST_FILL, ST_ADJOINT_FFT, ST_GATHER, ST_SORT_FINE
};
void timing_begin(int stage, hipStream_t s);
void timing_end(int stage, hipStream_t s);

struct StageTimer {
  int stage; hipStream_t s;
  StageTimer(int st, hipStream_t str) : stage(st), s(str) { timing_begin(stage, s); }
  ~StageTimer() { timing_end(stage, s); }
};

// ------------------------------------------------------------------------------------
// complex helpers
// ------------------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * (-i)
__device__ __forceinline__ float2 cmul_mi(float2 a) { return make_float2(a.y, -a.x); }

// forward DFTs (e^{-2 pi i r k / R}) in registers, natural-order output
template <int R> struct DFT;

template <> struct DFT<2> {
  static __device__ __forceinline__ void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <> struct DFT<3> {
  static __device__ __forceinline__ void run(float2* v) {
    const float s = 0.86602540378443864676f;  // sin(pi/3)
    float2 t = cadd(v[1], v[2]);
    float2 d = csub(v[1], v[2]);
    float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
    v[0] = cadd(v[0], t);
    // X1 = m - i s d ; X2 = m + i s d
    v[1] = make_float2(m.x + s * d.y, m.y - s * d.x);
    v[2] = make_float2(m.x - s * d.y, m.y + s * d.x);
  }
};

template <> struct DFT<4> {
  static __device__ __forceinline__ void run(float2* v) {
    float2 t0 = cadd(v[0], v[2]);
    float2 t1 = csub(v[0], v[2]);
    float2 t2 = cadd(v[1], v[3]);
    float2 t3 = cmul_mi(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

template <> struct DFT<8> {
  static __device__ __forceinline__ void run(float2* v) {
    const float c = 0.70710678118654752440f;  // sqrt(1/2)
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    DFT<4>::run(e);
    DFT<4>::run(o);
    // o[k] *= e^{-i pi k / 4}
    float2 o1 = make_float2(c * (o[1].x + o[1].y), c * (o[1].y - o[1].x));
    float2 o2 = cmul_mi(o[2]);
    float2 o3 = make_float2(c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

// ------------------------------------------------------------------------------------
// radix plan: as many 8s as possible, then 4, then 2, then 3
// ------------------------------------------------------------------------------------
__host__ __device__ constexpr int radix_next(int m) {
  return (m % 8 == 0) ? 8 : (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : (m % 3 == 0) ? 3 : 1;
}
__host__ __device__ constexpr int radix_at(int N, int s) {
  int m = N;
  for (int st = 0; st < 16; ++st) {
    int r = radix_next(m);
    if (m == 1 || r == 1) return 1;
    if (st == s) return r;
    m /= r;
  }
  return 1;
}
__host__ __device__ constexpr bool fft_size_ok(int N) {
  int m = N;
  for (int st = 0; st < 16 && m > 1; ++st) {
    int r = radix_next(m);
    if (r == 1) return false;
    m /= r;
  }
  return m == 1;
}

// ------------------------------------------------------------------------------------
// twiddle sources: a full table e^{-2 pi i q / M} (LDS or global), or its first quarter
// q < M / 4 (in LDS) with w^(a + b M/4) = w^a (-i)^b, exact up to the table's rounding
// ------------------------------------------------------------------------------------
__device__ __forceinline__ float2 tw_at(const float2* tw, int q) { return tw[q]; }

template <int M> struct TwQuarter {
  const float2* p;
};
template <int M> __device__ __forceinline__ float2 tw_at(TwQuarter<M> t, int q) {
  constexpr int Q = M / 4;
  static_assert(M % 4 == 0, "quarter tables need M divisible by 4");
  const int b = q / Q;
  const float2 v = t.p[q - b * Q];
  const float2 r = (b & 1) ? make_float2(v.y, -v.x) : v;
  return (b & 2) ? make_float2(-r.x, -r.y) : r;
}

// first half q < M / 2 (in LDS), with w^(q + M/2) = -w^q: half the table of the full
// one, and a negation instead of the quarter table's rotation logic
template <int M> struct TwHalf {
  const float2* p;
};
template <int M> __device__ __forceinline__ float2 tw_at(TwHalf<M> t, int q) {
  constexpr int H = M / 2;
  static_assert(M % 2 == 0, "half tables need M even");
  const bool b = q >= H;
  const float2 v = t.p[b ? q - H : q];
  return b ? make_float2(-v.x, -v.y) : v;
}

// ------------------------------------------------------------------------------------
// One Stockham stage over C independent length-N transforms held in LDS.
// Element (point p, column c) lives at buf[p * PS + c * CS].  The stage reads radix-R
// butterflies at stride N/R, applies twiddles w^(r k) with k = j mod NS, runs the
// R-point DFT and writes the outputs at (j / NS) * NS * R + k + r * NS (autosort).
// All reads finish (barrier) before any write, so one buffer suffices.
// tw: table of e^{-2 pi i q / (N * TWS)}, q < N * TWS (stride TWS selects length N).
// COLFAST: consecutive threads take consecutive columns (else consecutive butterflies).
// ------------------------------------------------------------------------------------
template <int N, int R, int NS, int C, int PS, int CS, int NT, bool COLFAST, int TWS, int SK, class TW>
__device__ __forceinline__ void fft_stage(float2* buf, TW tw) {
  // point p of column c: p * PS + c * CS, plus p >> SK with a skew (SK > 0), which spreads
  // the stride-R autosort writes over the LDS banks
  auto at = [](int p, int c) { return p * PS + (SK > 0 ? (p >> SK) : 0) + c * CS; };
  constexpr int B = N / R;
  constexpr int TASKS = B * C;
  constexpr int PER = (TASKS + NT - 1) / NT;
  const int tid = threadIdx.x;
  float2 v[PER][R];
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int task = tid + t * NT;
    if ((TASKS % NT) == 0 || task < TASKS) {
      const int j = COLFAST ? task / C : task % B;
      const int c = COLFAST ? task % C : task / B;
#pragma unroll
      for (int r = 0; r < R; ++r) v[t][r] = buf[at(j + r * B, c)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int task = tid + t * NT;
    if ((TASKS % NT) == 0 || task < TASKS) {
      const int j = COLFAST ? task / C : task % B;
      const int c = COLFAST ? task % C : task / B;
      const int k = j % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[t][r] = cmul(v[t][r], tw_at(tw, (r * k * (N / (NS * R))) * TWS));
      }
      DFT<R>::run(v[t]);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[at(base + r * NS, c)] = v[t][r];
    }
  }
  __syncthreads();
}

template <int N, int S, int NS, int C, int PS, int CS, int NT, bool COLFAST, int TWS, int SK>
struct FFTStages {
  template <class TW> static __device__ __forceinline__ void run(float2* buf, TW tw) {
    constexpr int R = radix_at(N, S);
    if constexpr (R > 1) {
      fft_stage<N, R, NS, C, PS, CS, NT, COLFAST, TWS, SK>(buf, tw);
      FFTStages<N, S + 1, NS * R, C, PS, CS, NT, COLFAST, TWS, SK>::run(buf, tw);
    }
  }
};

// forward length-N transforms of C columns in LDS (see fft_stage for the layout)
template <int N, int C, int PS, int CS, int NT, bool COLFAST, int TWS, int SK = 0, class TW>
__device__ __forceinline__ void lds_fft(float2* buf, TW tw) {
  static_assert(fft_size_ok(N), "unsupported FFT length");
  static_assert(SK == 0 || PS == 1, "skewed layouts take unit point stride");
  FFTStages<N, 0, 1, C, PS, CS, NT, COLFAST, TWS, SK>::run(buf, tw);
}

// Copies an R x W block of complex values (global row stride gs, LDS row stride PS) into
// LDS with every global load issued before the first LDS store, so each wave keeps all of
// its loads in flight instead of waiting on them one at a time.
template <int R, int W, int PS, int NT>
__device__ __forceinline__ void stage_tile(float2* __restrict__ buf, const float2* __restrict__ src, size_t gs) {
  constexpr int TOT = R * W, PER = (TOT + NT - 1) / NT;
  const int tid = threadIdx.x;
  float2 v[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = tid + k * NT;
    if (TOT % NT == 0 || i < TOT) v[k] = src[(size_t)(i / W) * gs + (i % W)];
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = tid + k * NT;
    if (TOT % NT == 0 || i < TOT) buf[(i / W) * PS + (i % W)] = v[k];
  }
}

// ------------------------------------------------------------------------------------
// misc
// ------------------------------------------------------------------------------------
__host__ __device__ constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v / 2); }
__host__ __device__ constexpr int pow2_floor(int v) { return v <= 1 ? 1 : 2 * pow2_floor(v / 2); }
// largest power of two dividing v
__host__ __device__ constexpr int pow2_div(int v) { return (v & 1) ? 1 : 2 * pow2_div(v / 2); }
__host__ __device__ constexpr int isqrt_c(int v) {
  int r = 0;
  while ((r + 1) * (r + 1) <= v) ++r;
  return r;
}
__host__ __device__ constexpr int cmin(int a, int b) { return a < b ? a : b; }
__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }

// Fixed-order block sum for NT a multiple of 64: xor-butterfly sums inside each wave, then
// wave 0 combines the per-wave totals the same way (two barriers instead of a log2(NT)
// tree).  red needs NT / 64 + 1 entries.  Every thread must call it; the result is in all
// threads, and every thread has passed both barriers before any reads it.
template <int NT>
__device__ __forceinline__ double block_sum_f64_waves(double v, double* red) {
  static_assert(NT % 64 == 0, "whole waves only");
  constexpr int NW = NT / 64;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x < 64) {
    double t = threadIdx.x < NW ? red[threadIdx.x] : 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
      if (off < NW * 2) t += __shfl_xor(t, off, 64);
    if (threadIdx.x == 0) red[NW] = t;
  }
  __syncthreads();
  return red[NW];
}

// The same for int64 (exact, so any order gives the same result).  red needs NT / 64 + 1
// entries; every thread must call it; the result is in all threads.
template <int NT>
__device__ __forceinline__ long long block_sum_i64_waves(long long v, long long* red) {
  static_assert(NT % 64 == 0, "whole waves only");
  constexpr int NW = NT / 64;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t = 0;
    for (int w = 0; w < NW; ++w) t += red[w];
    red[NW] = t;
  }
  __syncthreads();
  return red[NW];
}

__host__ __device__ constexpr int pow2_ceil(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Fixed-order (deterministic) block sum for any block size nt <= red's length.
// Every thread of the block must call it; ends with a barrier, result in all threads.
__device__ __forceinline__ double block_sum_f64(double v, double* red, int nt) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int s = pow2_ceil(nt) / 2; s > 0; s >>= 1) {
    if (tid < s && tid + s < nt) red[tid] += red[tid + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

}  // namespace pk
