// Fused mass assignment + z pass of the R2C FFT (the mesh is never written to HBM).
//
// One block per row-bundle tile (x = xt, ty rows in y, all z): the sorted particle records
// of the tile and of the neighbour buckets that spill into it are deposited into the ty
// rows held in LDS, the estimated mean density c = sum(w) / n^3 is subtracted, and each
// row is transformed along z (length-n/2 complex FFT + split, as the stand-alone z pass).
// The spectrum rows go to [kz][x][y] like the z pass, so the y and x passes follow
// unchanged.
//
// Contrast (tests/lognormal.py:79-80): delta = rho / rhobar - 1.  Subtracting any constant
// changes only the k = 0 mode, and rhobar = M / n^3 is only known once every tile has
// deposited, so the rows are transformed as rho - c and the per-bin sums are scaled by
// (n^3 / M)^2 afterwards; k = 0 is taken as exactly 0.  c removes the large k = 0 value
// from the float32 butterflies.
#include "pk_particles.h"

namespace pk {

// Row-bundle height of the default tile (the rule of make_tile_geom: the largest divisor of
// N that is <= 32 and whose rows fit the 64-bit fixed-point mesh deposit, so 16 at N = 768
// and 1024) and of the second compiled variant (half of it): the autotuner picks between
// them.
template <int N> struct FusedTY {
  static constexpr int D32 = N <= 32 ? N : (N % 32 == 0 ? 32 : (N % 24 == 0 ? 24 : 16));
  static constexpr int DEF = (size_t)D32 * N * 8 <= FIXED_LDS_MAX ? D32 : 16;
  static constexpr int ALT = (N >= 64 && DEF % 2 == 0 && N % (DEF / 2) == 0) ? DEF / 2 : DEF;
};

// The tile's cells are 64-bit fixed-point integers stored z-major, [N][TY] (cell (row, z) at
// z TY + row), so the lanes of the z transform's first stage, which take consecutive rows,
// read consecutive cells.  That stage converts them to float (rho - c) in registers and
// writes its outputs as the complex rows [TY][N/2 + 1] that the later stages use, in the
// same LDS (all of its reads finish before its writes).
template <int N, int TY_> struct FusedCfg {
  static constexpr int NC = N / 2;
  static constexpr int CS = NC + 1;                      // complex row stride in LDS
  static constexpr int RS = 2 * CS;                      // float row stride
  static constexpr int TY = TY_;
  static constexpr size_t LDS = cmax((int)((size_t)TY * N * 8), (int)((size_t)TY * CS * sizeof(float2)));
  static constexpr size_t STATIC_LDS = (size_t)N * 8 + 2048;   // twiddles, bucket tables, sums
  // blocks of this tile size a CU holds; with two or more, 512 threads each (registers)
  static constexpr int BPC = (int)((160 * 1024) / (LDS + STATIC_LDS));
  // at least 512 threads: small meshes have dense tiles, and the deposit of their records
  // needs the parallelism even where the z transform alone would use fewer threads
  static constexpr int NT0 = cmin(1024, cmax(cmin(512, TY * N), (TY * NC / 8) / 64 * 64));
  static constexpr int NT = BPC >= 2 ? cmin(NT0, 512) : NT0;
  static constexpr int PER = (TY * N + NT - 1) / NT;     // cells per thread in the conversion
  // waves per SIMD the LDS allows (capped at 4): the register budget is held to it
  static constexpr int WPE = cmax(1, cmin(4, cmax(1, BPC) * (NT / 64) / 4));
};

// Bucket descriptors of one tile in LDS: first record and exclusive prefix of sizes.
struct TileBuckets {
  int base[MAX_TILE_BUCKETS];
  int pre[MAX_TILE_BUCKETS + 1];
  int nbk;
};

// Deposits are rint(W w S) with S = 2^(30 - e), max |w| < 2^e and corner weights W <= 1, so
// each is below 2^30 and converts through int32; a cell of R records sums below R 2^30 and a
// tile's cells below 9 R 2^30, inside int64 for any catalogue of < 2^29 particles.  The
// quantum 2^(e - 30) must stay this fine: the rounding is white noise on the mesh, and it
// enters |delta_k|^2 through the cross term with the signal, i.e. as the square root of its
// power, which the TSC window's deconvolution at the corner of the spectrum exposes
// (a 2^-18 quantum measured 1e-4 there, 70 times the float32 FFT floor).
// ST: the stencil fixed at compile time (plain launches), or -1 for the run-time value of
// each job (batched launches mix stencils)
template <int N, int TY, int ST>
__device__ __forceinline__ void fused_body(const Job& j, int bx, int nbx, const float2* __restrict__ twg) {
  using Cf = FusedCfg<N, TY>;
  constexpr int NC = Cf::NC, CS = Cf::CS, C = Cf::TY, NT = Cf::NT;
  constexpr int U = 3;   // records per thread loaded ahead (2 and 4 measured slower)
  extern __shared__ float2 buf[];   // N x C int64 cells (z-major), then C rows of CS complex
  __shared__ float2 tw[N];   // full table: the split and the stages read one entry each
  __shared__ TileBuckets tb[2];
  __shared__ unsigned long long tsum[2];   // exact integer mass of the last two tiles
  __shared__ double wsq[NT / 64];           // per-wave sums of the tile's (rho - c)^2
  const ScatterArgs a = j.a;
  const TileGeom g = j.g;
  const void* sorted = j.sorted;
  const bool has_w = a.w != nullptr;
  const int stencil = ST >= 0 ? ST : a.stencil;
  const bool tsc = stencil == PK_STENCIL_TSC;
  const bool irregular = stencil == PK_STENCIL_CIC_REF;
  unsigned long long* acc = reinterpret_cast<unsigned long long*>(buf);
  const int tid = threadIdx.x;
  const int ntiles = g.ntiles;
  if (bx >= ntiles) return;   // whole block
  for (int i = tid; i < N; i += NT) tw[i] = twg[i];
  if (tid < 2) tsum[tid] = 0ull;
  const float scale = (float)j.aux[0];
  const double inv_scale = j.aux[1];
  int inv_exp;   // inv_scale = 2^inv_exp
  (void)frexp(inv_scale, &inv_exp);
  inv_exp -= 1;
  const float c0 = (float)j.stats[3];
  float2* spec = j.spec;
  const int* __restrict__ key_start = j.key_start;
  const int nbk = tile_bucket_count(g, N, tsc, irregular);   // the same for every tile
  // Bucket bounds of tile t: lane q < nbk loads its own key's [start, end) and does not
  // wait, so the loads of the next tile's bounds land while this tile deposits.
  auto issue = [&](int t, int& lo, int& hi) {
    if (tid < nbk) {
      const int key = tile_bucket_key(g, N, tsc, t / g.nyt, t % g.nyt, tid);
      lo = key_start[key];
      hi = key_start[key + 1];
    }
  };
  // ... then wave 0 scans the sizes into B (one barrier, which also ends the deposit)
  auto publish = [&](TileBuckets& B, int lo, int hi) {
    if (tid < 64) {
      int incl = tid < nbk ? hi - lo : 0;
#pragma unroll
      for (int off = 1; off < MAX_TILE_BUCKETS; off <<= 1) {
        const int v = __shfl_up(incl, off, 64);
        if (tid >= off) incl += v;
      }
      if (tid < nbk) {
        B.base[tid] = lo;
        B.pre[tid + 1] = incl;
      }
      if (tid == 0) {
        B.pre[0] = 0;
        B.nbk = nbk;
      }
    }
    __syncthreads();
  };
  // record idx of the tile buckets B (clamped to the last one, so the loads of a thread's
  // records need no branch and are all in flight together; callers test idx < total at the
  // deposit): a fixed four-step search for the bucket holding idx.  The irregular bucket is
  // the last one.
  static_assert(MAX_TILE_BUCKETS <= 16, "four search steps");
  auto record = [&](const TileBuckets& B, int idx) -> float4 {
    const int i = min(idx, B.pre[B.nbk] - 1);
    int b = 0;
#pragma unroll
    for (int step = 8; step >= 1; step >>= 1) {
      const int p = B.pre[b + step];   // b + step <= 15: inside pre[] even beyond nbk
      b = (b + step < B.nbk && p <= i) ? b + step : b;
    }
    return load_record_stream(sorted, (size_t)B.base[b] + (i - B.pre[b]), has_w);
  };
  auto regular = [&](const TileBuckets& B, int idx) -> bool { return !irregular || idx < B.pre[B.nbk - 1]; };
  // Software pipeline over this block's tiles: the next tile's bucket bounds load during the
  // current deposit, and its first U * NT records are loaded into registers before the
  // current tile's FFT, so their latency overlaps the transform and the writes.
  int lo = 0, hi = 0;
  issue(bx, lo, hi);
  publish(tb[0], lo, hi);
  float4 rp[U];
  auto prefetch = [&](const TileBuckets& B) {
    if (B.pre[B.nbk] > 0) {   // block-uniform
#pragma unroll
      for (int u = 0; u < U; ++u) rp[u] = record(B, tid + u * NT);
    }
  };
  prefetch(tb[0]);
  int cur = 0;
  int t_prev = -1;
  for (int t = bx; t < ntiles; t += nbx) {
    const int xt = t / g.nyt, yt = t % g.nyt;
    const int y0 = yt * C;
    const int tn = t + nbx;
    const TileBuckets& B = tb[cur];
    if (tn < ntiles) issue(tn, lo, hi);
    {
      uint4* z4 = reinterpret_cast<uint4*>(acc);   // C N 8 bytes: whole 16-byte words
      for (int i = tid; i < C * N / 2; i += NT) z4[i] = make_uint4(0u, 0u, 0u, 0u);
    }
    __syncthreads();
    // the previous tile's exact mass (every wave added its share before the barrier above)
    if (tid == 0 && t_prev >= 0) {
      j.tile_mass[t_prev] = (double)(long long)tsum[cur ^ 1] * inv_scale;
      tsum[cur ^ 1] = 0ull;
    }
    // the record's weight carries the power-of-two scale, so each corner is (W w) S exactly
    // and costs only the round and the int32 conversion
    auto add = [&](int iy, int iz, float wt) {
      const int ly = iy - y0;
      if ((unsigned)ly < (unsigned)C) atomicAdd(&acc[iz * C + ly], to_fixed(wt));
    };
    auto dep = [&](const float4& r, bool reg) {
      const float ws = record_weight(r, has_w) * scale;
      if (stencil == PK_STENCIL_CIC_REF && reg) ref_corners_on_x_regular(a, r.x, r.y, r.z, ws, xt, add);
      else for_each_corner_on_x<ST>(a, r.x, r.y, r.z, ws, xt, add);
    };
    const int total = B.pre[B.nbk];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (tid + u * NT < total) dep(rp[u], regular(B, tid + u * NT));
    for (int idx0 = tid + U * NT; idx0 < total; idx0 += U * NT) {
      float4 c[U];
#pragma unroll
      for (int u = 0; u < U; ++u) c[u] = record(B, idx0 + u * NT);
      __builtin_amdgcn_sched_barrier(0);   // every load issued before the first deposit
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (idx0 + u * NT < total) dep(c[u], regular(B, idx0 + u * NT));
    }
    if (tn < ntiles)   // block-uniform; the barrier also closes the deposit
      publish(tb[cur ^ 1], lo, hi);
    else
      __syncthreads();
    // The z transform (pairs of reals as NC complex, then the split).  Its first radix-R0
    // stage reads the integer cells, converts rho - c in registers and sums the tile's
    // integer mass (exact in 64 bits, one LDS add per wave); consecutive lanes take
    // consecutive rows, so the butterflies' LDS accesses are free of conflicts.
    {
      constexpr int R0 = radix_at(NC, 0);
      constexpr int B0 = NC / R0;
      constexpr int TASKS = B0 * C;
      constexpr int PER0 = (TASKS + NT - 1) / NT;
      float2 v[PER0][R0];
      long long isum = 0;
      float sqf = 0.0f;   // (rho - c)^2 of this thread's cells, in a fixed order
#pragma unroll
      for (int u = 0; u < PER0; ++u) {
        const int task = tid + u * NT;
        if (TASKS % NT == 0 || task < TASKS) {
          const int jj = task / C, c = task % C;
          long long q[2 * R0];
          bool fit = true;
#pragma unroll
          for (int r = 0; r < R0; ++r) {
            const int p = jj + r * B0;   // complex point p = cells z = 2p, 2p + 1
            q[2 * r] = (long long)acc[(2 * p) * C + c];
            q[2 * r + 1] = (long long)acc[(2 * p + 1) * C + c];
            isum += q[2 * r] + q[2 * r + 1];
            fit = fit && (long long)(int)q[2 * r] == q[2 * r] && (long long)(int)q[2 * r + 1] == q[2 * r + 1];
          }
          // rho - c: float(q) 2^inv_exp equals float(q inv_scale) (scaling by a power of two
          // commutes with the rounding), so when every cell of the wave is below 2^31 in
          // magnitude (nearly always) they convert through int32, not int64 -> double -> float
          if (__builtin_amdgcn_ballot_w64(!fit) == 0) {
#pragma unroll
            for (int r = 0; r < R0; ++r)
              v[u][r] = make_float2(ldexpf((float)(int)q[2 * r], inv_exp) - c0,
                                    ldexpf((float)(int)q[2 * r + 1], inv_exp) - c0);
          } else {
#pragma unroll
            for (int r = 0; r < R0; ++r)
              v[u][r] = make_float2((float)((double)q[2 * r] * inv_scale) - c0,
                                    (float)((double)q[2 * r + 1] * inv_scale) - c0);
          }
#pragma unroll
          for (int r = 0; r < R0; ++r) sqf = fmaf(v[u][r].x, v[u][r].x, fmaf(v[u][r].y, v[u][r].y, sqf));
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) isum += __shfl_xor(isum, off, 64);
      if ((tid & 63) == 0) atomicAdd(&tsum[cur], (unsigned long long)isum);
      {   // the tile's sum of squares (Parseval bound of the binned power), fixed order
        double sq = (double)sqf;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
        if ((tid & 63) == 0) wsq[tid >> 6] = sq;
      }
      if (tn < ntiles) prefetch(tb[cur ^ 1]);
      __syncthreads();   // every cell read before the complex rows overwrite them
      if (tid == 0 && j.tile_sq != nullptr) {
        double sq = 0.0;
        for (int w = 0; w < NT / 64; ++w) sq += wsq[w];
        j.tile_sq[t] = sq;
      }
#pragma unroll
      for (int u = 0; u < PER0; ++u) {
        const int task = tid + u * NT;
        if (TASKS % NT == 0 || task < TASKS) {
          const int jj = task / C, c = task % C;
          DFT<R0>::run(v[u]);
#pragma unroll
          for (int r = 0; r < R0; ++r) buf[(jj * R0 + r) + c * CS] = v[u][r];
        }
      }
      __syncthreads();
      FFTStages<NC, 1, R0, C, 1, CS, NT, true, 2, 0>::run(buf, tw);
    }
    for (int i = tid; i < (NC + 1) * C; i += NT) {
      const int c = i % C, k = i / C;
      const float2 zk = buf[c * CS + (k == NC ? 0 : k)];
      const float2 zr = buf[c * CS + (k == 0 ? 0 : NC - k)];
      const float2 zc = make_float2(zr.x, -zr.y);
      const float2 fe = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y + zc.y));
      const float2 dd = csub(zk, zc);
      const float2 fo = make_float2(0.5f * dd.y, -0.5f * dd.x);
      spec[((size_t)k * N + xt) * N + (size_t)y0 + c] = cadd(fe, cmul(tw[k], fo));
    }
    __syncthreads();   // the split has read the rows before the next tile clears them
    t_prev = t;
    cur ^= 1;
  }
  if (tid == 0 && t_prev >= 0) j.tile_mass[t_prev] = (double)(long long)tsum[cur ^ 1] * inv_scale;
}

template <int N, int TY, int ST>
__global__ void __launch_bounds__((FusedCfg<N, TY>::NT)) __attribute__((amdgpu_waves_per_eu(FusedCfg<N, TY>::WPE)))
k_fused_deposit_zpass(Job j, const float2* __restrict__ tw) {
  fused_body<N, TY, ST>(j, blockIdx.x, gridDim.x, tw);
}

template <int N, int TY>
__global__ void __launch_bounds__((FusedCfg<N, TY>::NT)) __attribute__((amdgpu_waves_per_eu(FusedCfg<N, TY>::WPE)))
k_fused_deposit_zpass_jobs(const Job* __restrict__ jobs, const float2* __restrict__ tw) {
  const Job& j = jobs[blockIdx.y];
  if (!j.active) return;
  fused_body<N, TY, -1>(j, blockIdx.x, gridDim.x, tw);
}

template <int N, int TY> struct FusedLaunch {
  using Cf = FusedCfg<N, TY>;
  static bool ok(const TileGeom& g) { return g.ty == TY && Cf::LDS + Cf::STATIC_LDS <= 160 * 1024; }
  template <typename K>
  static int attr(K k) {
    if (Cf::LDS <= 65536) return PK_OK;
    return check_hip(hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)Cf::LDS),
                     "LDS attribute");
  }
  // persistent grid: resident blocks of the whole GPU (LDS-limited), at most one per tile
  static int grid() {
    static int g = 0;
    if (!g) {
      int dev = 0, occ = 0;
      (void)hipGetDevice(&dev);
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_fused_deposit_zpass<N, TY, PK_STENCIL_CIC>, Cf::NT,
                                                       Cf::LDS) !=
              hipSuccess ||
          occ < 1)
        occ = 1;
      g = num_cus(dev) * occ;
    }
    return g;
  }
  template <int ST>
  static int run_st(const Job& j, const float2* tw, hipStream_t s) {
    static bool done = false;
    if (!done) {
      int e = attr(k_fused_deposit_zpass<N, TY, ST>);
      if (e) return e;
      done = true;
    }
    const int gx = j.g.ntiles < grid() ? j.g.ntiles : grid();
    StageTimer t(ST_DEPOSIT, s);
    hipLaunchKernelGGL((k_fused_deposit_zpass<N, TY, ST>), dim3(gx), dim3(Cf::NT), Cf::LDS, s, j, tw);
    return check_hip(hipGetLastError(), "fused deposit + z pass launch");
  }
  static int run(const Job& j, const float2* tw, hipStream_t s) {
    if (j.a.stencil == PK_STENCIL_CIC_REF) return run_st<PK_STENCIL_CIC_REF>(j, tw, s);
    if (j.a.stencil == PK_STENCIL_CIC) return run_st<PK_STENCIL_CIC>(j, tw, s);
    return run_st<PK_STENCIL_TSC>(j, tw, s);
  }
  static int run_jobs(const Job* jobs, int njobs, int ntiles, const float2* tw, hipStream_t s) {
    static bool done = false;
    if (!done) {
      int e = attr(k_fused_deposit_zpass_jobs<N, TY>);
      if (e) return e;
      done = true;
    }
    // one block per tile: jobs of very different sizes share the launch, and dynamic block
    // scheduling balances them better than a fixed share of a persistent grid per job
    const int gx = ntiles;
    StageTimer t(ST_DEPOSIT, s);
    hipLaunchKernelGGL((k_fused_deposit_zpass_jobs<N, TY>), dim3(gx, njobs), dim3(Cf::NT), Cf::LDS, s, jobs, tw);
    return check_hip(hipGetLastError(), "batched fused deposit + z pass launch");
  }
};

#define PK_FOR_EACH_FUSED_N(X) \
  X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)

bool fused_supported(int n, const TileGeom& g) {
  switch (n) {
#define PK_CASE(V)                                                                    \
  case V:                                                                             \
    return FusedLaunch<V, FusedTY<V>::DEF>::ok(g) || FusedLaunch<V, FusedTY<V>::ALT>::ok(g);
    PK_FOR_EACH_FUSED_N(PK_CASE)
#undef PK_CASE
    default: return false;
  }
}

int fused_deposit_zpass(int n, const Job& j, const float2* tw, hipStream_t s) {
  switch (n) {
#define PK_CASE(V)                                                                         \
  case V:                                                                                  \
    if (j.g.ty == FusedTY<V>::DEF) return FusedLaunch<V, FusedTY<V>::DEF>::run(j, tw, s);  \
    if (j.g.ty == FusedTY<V>::ALT) return FusedLaunch<V, FusedTY<V>::ALT>::run(j, tw, s);  \
    break;
    PK_FOR_EACH_FUSED_N(PK_CASE)
#undef PK_CASE
  }
  return set_error(PK_E_INVALID, "fused path: no kernel for n=%d, %d rows per tile", n, j.g.ty);
}

int fused_deposit_zpass_jobs(int n, int ty, const Job* jobs, int njobs, int ntiles, const float2* tw, hipStream_t s) {
  switch (n) {
#define PK_CASE(V)                                                                                           \
  case V:                                                                                                    \
    if (ty == FusedTY<V>::DEF) return FusedLaunch<V, FusedTY<V>::DEF>::run_jobs(jobs, njobs, ntiles, tw, s);  \
    if (ty == FusedTY<V>::ALT) return FusedLaunch<V, FusedTY<V>::ALT>::run_jobs(jobs, njobs, ntiles, tw, s);  \
    break;
    PK_FOR_EACH_FUSED_N(PK_CASE)
#undef PK_CASE
  }
  return set_error(PK_E_INVALID, "fused path: no kernel for n=%d, %d rows per tile", n, ty);
}

}  // namespace pk
