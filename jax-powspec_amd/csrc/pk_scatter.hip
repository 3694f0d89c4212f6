// Particle -> mesh mass assignment: reference stencil (src/mas.py:88-153 and :5-87),
// textbook CIC and TSC.
//
// Owner-computes tiled deposit:
//   1. count:  each particle gets a key = (home tile, boundary class); home = the base cell
//              (reference stencil / CIC) or the nearest grid point (TSC).  The boundary
//              class records on which faces of its tile the stencil spills into a
//              neighbouring tile.  Particles the tiles cannot take (reference stencil with an
//              out-of-range base index) get the last key ("irregular").
//   2. scan:   exclusive prefix sum of the key counts.
//   3. place:  particles are written, key-sorted, as float4(xpos, ypos, zpos, w).
//   4. deposit: one block per tile loads its cells into LDS (zeros or mesh_in), walks its
//              own buckets plus the neighbour buckets whose class spills into it, adds
//              every stencil corner that falls inside the tile with LDS float atomics and
//              writes the tile once with coalesced stores: each mesh cell is written by
//              exactly one block, with no global atomics and no separate zero fill.
//   5. atomic: the (normally empty) irregular bucket is added with global atomics.
// A direct mode (tile[0] == 0) skips 1-4 and adds every particle with global atomics.
//
// The position -> (cell, fraction) arithmetic is float32 exactly as the reference
// evaluates it on float32 inputs: pos = (x - xmin) * inv (src/mas.py:103-105),
// ix = int32(pos) (truncation, :107-109), dd = pos - ix (:111-113), md = 1 - dd (:138-140).
#include "pk_common.h"
#include "pk_internal.h"

namespace pk {

constexpr float POS_LIMIT = 1073741824.0f;  // |pos| >= 2^30 or non-finite: particle dropped

__device__ __forceinline__ int floor_mod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

__device__ __forceinline__ void load_particle(const ScatterArgs& a, int64_t i, float& px, float& py, float& pz,
                                              float& w) {
  px = (a.x[i * a.sx] - a.xmin) * a.inv;
  py = (a.y[i * a.sy] - a.ymin) * a.inv;
  pz = (a.z[i * a.sz] - a.zmin) * a.inv;
  w = a.w ? a.w[i * a.sw] : 1.0f;
}

__device__ __forceinline__ bool pos_ok(float px, float py, float pz) {
  return fabsf(px) < POS_LIMIT && fabsf(py) < POS_LIMIT && fabsf(pz) < POS_LIMIT;  // false for NaN
}

// one axis of the reference stencil: base index b, +1 index p (wrap rule of `mode`),
// validity after JAX index normalisation, and fractions md, dd
struct RefAxis {
  int b, p;
  bool vb, vp;
  float md, dd;
};

__device__ __forceinline__ RefAxis ref_axis(float pos, int n, int mode) {
  RefAxis r;
  const int i = (int)pos;  // truncation toward zero
  float d = pos - (float)i;
  r.md = 1.0f - d;
  int p = i + 1;
  if (mode == PK_WRAP_VEC_TRUE) {
    p = floor_mod(p + n, n);               // src/mas.py:121
  } else if (mode == PK_WRAP_VEC_FALSE) {
    p = (p >= n) ? 0 : p;                  // src/mas.py:125 (ddx test at :126 never fires)
  } else if (mode == PK_WRAP_SCAN_TRUE) {
    p = (p >= n) ? p - n : p;              // src/mas.py:24
  } else {                                 // PK_WRAP_SCAN_FALSE, src/mas.py:34
    if (p >= n) {
      p = 0;
      d = 0.0f;
    }
  }
  r.dd = d;
  r.b = i < 0 ? i + n : i;
  r.vb = (r.b >= 0) && (r.b < n);
  r.p = p < 0 ? p + n : p;
  r.vp = (r.p >= 0) && (r.p < n);
  return r;
}

// home cell per axis (regular particles only)
__device__ __forceinline__ int home_cell(float pos, int n, int stencil) {
  if (stencil == PK_STENCIL_TSC) return floor_mod((int)floorf(pos + 0.5f), n);
  if (stencil == PK_STENCIL_CIC) return floor_mod((int)floorf(pos), n);
  const int i = (int)pos;
  return i < 0 ? i + n : i;
}

// key of a particle, -1 if dropped
__device__ __forceinline__ int particle_key(const ScatterArgs& a, const TileGeom& g, float px, float py, float pz) {
  if (!pos_ok(px, py, pz)) return -1;
  const int n = a.n;
  if (a.stencil == PK_STENCIL_CIC_REF) {
    const int ix = (int)px, iy = (int)py, iz = (int)pz;
    const int bx = ix < 0 ? ix + n : ix, by = iy < 0 ? iy + n : iy, bz = iz < 0 ? iz + n : iz;
    if (bx < 0 || bx >= n || by < 0 || by >= n || bz < 0 || bz >= n) return g.nkeys - 1;
  }
  const float pos[3] = {px, py, pz};
  int tile = 0, cls = 0, mul = 1;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const int c = home_cell(pos[ax], n, a.stencil);
    const int t = c / g.td[ax];
    const int local = c - t * g.td[ax];
    const int ext = min(g.td[ax], n - t * g.td[ax]);
    tile = tile * g.nt[ax] + t;
    int v;
    if (a.stencil == PK_STENCIL_TSC) {
      v = (local == 0) ? 1 : ((local == ext - 1) ? 2 : 0);
      cls += v * mul;
      mul *= 3;
    } else {
      v = (local == ext - 1) ? 1 : 0;
      cls += v * mul;
      mul *= 2;
    }
  }
  return tile * g.ncls + cls;
}

// Visit every stencil corner of one particle: f(ix, iy, iz, weight); indexes are in
// [0, n) for regular particles.  For the reference stencil `all_checks` applies the
// per-corner JAX drop rule (used by the atomic path for irregular particles).
template <typename F>
__device__ __forceinline__ void for_each_corner(const ScatterArgs& a, float px, float py, float pz, float w, F&& f) {
  const int n = a.n;
  if (a.stencil == PK_STENCIL_CIC_REF) {
    const RefAxis X = ref_axis(px, n, a.wrap_mode), Y = ref_axis(py, n, a.wrap_mode), Z = ref_axis(pz, n, a.wrap_mode);
    // src/mas.py:142-151, products evaluated left to right as in the reference
    if (X.vb && Y.vb && Z.vb) f(X.b, Y.b, Z.b, X.md * Y.md * Z.md * w);
    if (X.vp && Y.vb && Z.vb) f(X.p, Y.b, Z.b, X.dd * Y.md * Z.md * w);
    if (X.vb && Y.vp && Z.vb) f(X.b, Y.p, Z.b, X.md * Y.dd * Z.md * w);
    if (X.vb && Y.vb && Z.vp) f(X.b, Y.b, Z.p, X.md * Y.md * Z.dd * w);
    if (X.vp && Y.vp && Z.vb) f(X.p, Y.p, Z.b, X.dd * Y.dd * Z.md * w);
    if (X.vp && Y.vb && Z.vp) f(X.p, Y.b, Z.p, X.dd * Y.md * Z.dd * w);
    if (X.vb && Y.vp && Z.vp) f(X.b, Y.p, Z.p, X.md * Y.md * Z.dd * w);  // src/mas.py:149 weight
    if (X.vp && Y.vp && Z.vp) f(X.p, Y.p, Z.p, X.dd * Y.dd * Z.dd * w);
  } else if (a.stencil == PK_STENCIL_CIC) {
    const float pos[3] = {px, py, pz};
    int c0[3], c1[3];
    float w0[3], w1[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const float f0 = floorf(pos[ax]);
      const float d = pos[ax] - f0;
      c0[ax] = floor_mod((int)f0, n);
      c1[ax] = c0[ax] + 1 == n ? 0 : c0[ax] + 1;
      w0[ax] = 1.0f - d;
      w1[ax] = d;
    }
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int sx = s >> 2, sy = (s >> 1) & 1, sz = s & 1;
      f(sx ? c1[0] : c0[0], sy ? c1[1] : c0[1], sz ? c1[2] : c0[2],
        (sx ? w1[0] : w0[0]) * (sy ? w1[1] : w0[1]) * (sz ? w1[2] : w0[2]) * w);
    }
  } else {  // TSC
    const float pos[3] = {px, py, pz};
    int c[3][3];
    float wt[3][3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const float f0 = floorf(pos[ax] + 0.5f);
      const float d = pos[ax] - f0;
      const int i0 = floor_mod((int)f0, n);
      c[ax][0] = i0 == 0 ? n - 1 : i0 - 1;
      c[ax][1] = i0;
      c[ax][2] = i0 + 1 == n ? 0 : i0 + 1;
      const float hm = 0.5f - d, hp = 0.5f + d;
      wt[ax][0] = 0.5f * hm * hm;
      wt[ax][1] = 0.75f - d * d;
      wt[ax][2] = 0.5f * hp * hp;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) f(c[0][i], c[1][j], c[2][k], wt[0][i] * wt[1][j] * wt[2][k] * w);
  }
}

// ------------------------------------------------------------------------------------
// small utility kernels (fills instead of hipMemset so a captured graph has no memset node)
// ------------------------------------------------------------------------------------
__global__ void k_fill_i32(int* __restrict__ p, int64_t n, int v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f64(double* __restrict__ p, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_copy_or_zero(const float* src, float* dst, int64_t n) {  // src may alias dst
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src ? src[i] : 0.0f;
}

// ------------------------------------------------------------------------------------
// 1. count
// ------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_sort_count(ScatterArgs a, TileGeom g, int* __restrict__ counts) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.np; i += (int64_t)gridDim.x * blockDim.x) {
    float px, py, pz, w;
    load_particle(a, i, px, py, pz, w);
    const int key = particle_key(a, g, px, py, pz);
    if (key >= 0) atomicAdd(&counts[key], 1);
  }
}

// ------------------------------------------------------------------------------------
// 2. exclusive scan (3 kernels; up to 4096 * 4096 keys)
// ------------------------------------------------------------------------------------
constexpr int SCAN_NT = 1024;
constexpr int SCAN_PER = 4;
constexpr int SCAN_BLOCK = SCAN_NT * SCAN_PER;

__device__ int block_exclusive_scan(int v, int* sh, int& total) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int off = 1; off < SCAN_NT; off <<= 1) {
    const int t = tid >= off ? sh[tid - off] : 0;
    __syncthreads();
    sh[tid] += t;
    __syncthreads();
  }
  total = sh[SCAN_NT - 1];
  const int incl = sh[tid];
  __syncthreads();
  return incl - v;
}

__global__ void __launch_bounds__(SCAN_NT) k_scan_local(const int* __restrict__ in, int* __restrict__ out,
                                                        int* __restrict__ bsum, int n) {
  __shared__ int sh[SCAN_NT];
  const int base = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_PER;
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
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ void __launch_bounds__(SCAN_NT) k_scan_top(int* __restrict__ bsum, int nblk, int* __restrict__ total_out) {
  __shared__ int sh[SCAN_NT];
  const int base = threadIdx.x * SCAN_PER;
  int v[SCAN_PER];
  int s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k) {
    v[k] = (base + k < nblk) ? bsum[base + k] : 0;
    s += v[k];
  }
  int total;
  int run = block_exclusive_scan(s, sh, total);
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k) {
    if (base + k < nblk) bsum[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) *total_out = total;
}

__global__ void __launch_bounds__(SCAN_NT) k_scan_add(int* __restrict__ out, int* __restrict__ cursor,
                                                      const int* __restrict__ bsum, int n) {
  const int base = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_PER;
  const int add = bsum[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k) {
    if (base + k < n) {
      const int o = out[base + k] + add;
      out[base + k] = o;
      cursor[base + k] = o;
    }
  }
}

// ------------------------------------------------------------------------------------
// 3. place (key-sorted float4 records)
// ------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_sort_place(ScatterArgs a, TileGeom g, int* __restrict__ cursor,
                                                    float4* __restrict__ sorted) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.np; i += (int64_t)gridDim.x * blockDim.x) {
    float px, py, pz, w;
    load_particle(a, i, px, py, pz, w);
    const int key = particle_key(a, g, px, py, pz);
    if (key >= 0) {
      const int at = atomicAdd(&cursor[key], 1);
      sorted[at] = make_float4(px, py, pz, w);
    }
  }
}

// ------------------------------------------------------------------------------------
// 4. tile deposit (one block per tile)
// ------------------------------------------------------------------------------------
constexpr int DEP_NT = 256;
constexpr int MAX_BUCKETS = 125;

__device__ double block_sum_f64(double v, double* sh) {
  // fixed-order tree: deterministic
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int off = DEP_NT / 2; off > 0; off >>= 1) {
    if (tid < off) sh[tid] += sh[tid + off];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(DEP_NT) k_deposit_tiles(ScatterArgs a, TileGeom g, const int* __restrict__ offsets,
                                                          const float4* __restrict__ sorted,
                                                          const float* mesh_in, float* mesh_out,
                                                          double* __restrict__ tile_mass) {
  extern __shared__ float tilebuf[];
  __shared__ int s_pc[3][5], s_pk[3][5], s_np[3];
  __shared__ int s_base[MAX_BUCKETS], s_pre[MAX_BUCKETS + 1];
  __shared__ int s_nbk;
  __shared__ double s_red[DEP_NT];
  const int tid = threadIdx.x;
  const int n = a.n;
  const int tile = blockIdx.x;
  int tc[3];
  tc[2] = tile % g.nt[2];
  tc[1] = (tile / g.nt[2]) % g.nt[1];
  tc[0] = tile / (g.nt[2] * g.nt[1]);
  int org[3], ext[3];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    org[ax] = tc[ax] * g.td[ax];
    ext[ax] = min(g.td[ax], n - org[ax]);
  }
  const int tdy = g.td[1], tdz = g.td[2];
  const int cells = g.td[0] * tdy * tdz;
  // init the tile (cells outside ext are never touched); mesh_in may alias mesh_out, so it
  // is read only here, before this block writes any of its cells
  double minit = 0.0;
  for (int i = tid; i < cells; i += DEP_NT) {
    const int lz = i % tdz, ly = (i / tdz) % tdy, lx = i / (tdz * tdy);
    float v = 0.0f;
    if (mesh_in && lx < ext[0] && ly < ext[1] && lz < ext[2])
      v = mesh_in[((size_t)(org[0] + lx) * n + (org[1] + ly)) * n + (org[2] + lz)];
    tilebuf[i] = v;
    minit += (double)v;
  }
  // per-axis (tile coordinate, class) pairs whose particles reach this tile
  if (tid == 0) {
    const bool tsc = a.stencil == PK_STENCIL_TSC;
    int total = 1;
    for (int ax = 0; ax < 3; ++ax) {
      const int t = tc[ax], nt = g.nt[ax];
      int k = 0;
      if (tsc) {
        for (int v = 0; v < 3; ++v) { s_pc[ax][k] = t; s_pk[ax][k] = v; ++k; }
        if (nt >= 3) {
          s_pc[ax][k] = (t + nt - 1) % nt; s_pk[ax][k] = 2; ++k;   // lower tile, last layer
          s_pc[ax][k] = (t + 1) % nt;      s_pk[ax][k] = 1; ++k;   // upper tile, first layer
        } else if (nt == 2) {
          s_pc[ax][k] = 1 - t; s_pk[ax][k] = 1; ++k;
          s_pc[ax][k] = 1 - t; s_pk[ax][k] = 2; ++k;
        }
      } else {
        for (int v = 0; v < 2; ++v) { s_pc[ax][k] = t; s_pk[ax][k] = v; ++k; }
        if (nt >= 2) { s_pc[ax][k] = (t + nt - 1) % nt; s_pk[ax][k] = 1; ++k; }
      }
      s_np[ax] = k;
      total *= k;
    }
    s_nbk = total;
  }
  __syncthreads();
  const int nbk = s_nbk;
  int mysize = 0;
  if (tid < nbk) {
    const int i2 = tid % s_np[2], i1 = (tid / s_np[2]) % s_np[1], i0 = tid / (s_np[2] * s_np[1]);
    const int nt1 = g.nt[1], nt2 = g.nt[2];
    const int tl = (s_pc[0][i0] * nt1 + s_pc[1][i1]) * nt2 + s_pc[2][i2];
    const bool tsc = a.stencil == PK_STENCIL_TSC;
    const int cls = tsc ? s_pk[0][i0] + 3 * s_pk[1][i1] + 9 * s_pk[2][i2] : s_pk[0][i0] + 2 * s_pk[1][i1] + 4 * s_pk[2][i2];
    const int key = tl * g.ncls + cls;
    const int lo = offsets[key];
    mysize = offsets[key + 1] - lo;
    s_base[tid] = lo;
  }
  __syncthreads();
  // bucket sizes -> exclusive prefix (<= 125 entries, one thread)
  if (tid < nbk) s_pre[tid] = mysize;
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int b = 0; b < nbk; ++b) {
      const int sz = s_pre[b];
      s_pre[b] = run;
      run += sz;
    }
    s_pre[nbk] = run;
  }
  __syncthreads();
  const int total = s_pre[nbk];
  for (int idx = tid; idx < total; idx += DEP_NT) {
    int lo = 0, hi = nbk - 1;  // last b with s_pre[b] <= idx
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_pre[mid] <= idx) lo = mid; else hi = mid - 1;
    }
    const float4 r = sorted[s_base[lo] + (idx - s_pre[lo])];
    for_each_corner(a, r.x, r.y, r.z, r.w, [&](int ix, int iy, int iz, float wt) {
      const int lx = ix - org[0], ly = iy - org[1], lz = iz - org[2];
      if ((unsigned)lx < (unsigned)ext[0] && (unsigned)ly < (unsigned)ext[1] && (unsigned)lz < (unsigned)ext[2])
        atomicAdd(&tilebuf[(lx * tdy + ly) * tdz + lz], wt);
    });
  }
  __syncthreads();
  double msum = 0.0;
  for (int i = tid; i < cells; i += DEP_NT) {
    const int lz = i % tdz, ly = (i / tdz) % tdy, lx = i / (tdz * tdy);
    if (lx < ext[0] && ly < ext[1] && lz < ext[2]) {
      const size_t gi = ((size_t)(org[0] + lx) * n + (org[1] + ly)) * n + (org[2] + lz);
      const float v = tilebuf[i];
      msum += (double)v;
      mesh_out[gi] = v;
    }
  }
  const double m = block_sum_f64(msum - minit, s_red);
  if (tid == 0) tile_mass[tile] = m;
}

// ------------------------------------------------------------------------------------
// 5. global-atomic deposit: all particles (direct mode) or the irregular bucket
// ------------------------------------------------------------------------------------
__global__ void __launch_bounds__(DEP_NT) k_atomic_deposit(ScatterArgs a, const float4* __restrict__ sorted,
                                                           const int* __restrict__ range, float* __restrict__ mesh,
                                                           double* __restrict__ stats) {
  __shared__ double s_red[DEP_NT];
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
      const float4 r = sorted[i];
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
  const double m = block_sum_f64(msum, s_red);
  if (threadIdx.x == 0 && m != 0.0) atomicAdd(&stats[2], m);
}

// ------------------------------------------------------------------------------------
// mass finalize: stats[0] = M (fixed-order sum of tile masses + atomic part), stats[1] = n^3 / M
// ------------------------------------------------------------------------------------
__global__ void __launch_bounds__(DEP_NT) k_mass_finalize(const double* __restrict__ tile_mass, int ntiles,
                                                          double* __restrict__ stats, double ncells) {
  __shared__ double s_red[DEP_NT];
  double s = 0.0;
  for (int i = threadIdx.x; i < ntiles; i += DEP_NT) s += tile_mass[i];
  const double m = block_sum_f64(s, s_red) + stats[2];
  if (threadIdx.x == 0) {
    stats[0] = m;
    stats[1] = ncells / m;
  }
}

// ------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------
int make_tile_geom(int n, int stencil, const int* tile, TileGeom* g, bool* direct) {
  *direct = false;
  int td[3] = {16, 16, 32};
  if (tile) {
    if (tile[0] == 0) {
      *direct = true;
      *g = TileGeom{};
      return PK_OK;
    }
    td[0] = tile[0];
    td[1] = tile[1];
    td[2] = tile[2];
  }
  long cells = 1;
  for (int ax = 0; ax < 3; ++ax) {
    if (td[ax] > n) td[ax] = n;
    if (td[ax] < 2) return set_error(PK_E_INVALID, "tile dim %d < 2", td[ax]);
    const int rem = n % td[ax];
    if (rem == 1) return set_error(PK_E_INVALID, "tile dim %d leaves a 1-cell tile for n=%d", td[ax], n);
    g->td[ax] = td[ax];
    g->nt[ax] = (n + td[ax] - 1) / td[ax];
    cells *= td[ax];
  }
  if (cells * 4 > 65536) return set_error(PK_E_INVALID, "tile of %ld cells exceeds 64 KiB of LDS", cells);
  g->ncls = stencil == PK_STENCIL_TSC ? 27 : 8;
  g->ntiles = g->nt[0] * g->nt[1] * g->nt[2];
  const long nkeys = (long)g->ntiles * g->ncls + 1;
  if (nkeys > (long)SCAN_BLOCK * SCAN_BLOCK) return set_error(PK_E_INVALID, "too many tile buckets (%ld)", nkeys);
  g->nkeys = (int)nkeys;
  return PK_OK;
}

namespace {
struct ScatterWS {
  int* counts;
  int* offsets;
  int* cursor;
  int* bsum;
  float4* sorted;
  double* tile_mass;
  double* stats;
  size_t bytes;
};

ScatterWS carve(const TileGeom& g, bool direct, int64_t np, void* base) {
  ScatterWS w{};
  size_t off = 0;
  char* b = static_cast<char*>(base);
  auto take = [&](size_t bytes) -> void* {
    void* p = b ? b + off : nullptr;
    off += align_up(bytes);
    return p;
  };
  w.stats = static_cast<double*>(take(PK_STATS_LEN * sizeof(double)));
  if (!direct) {
    w.counts = static_cast<int*>(take((size_t)g.nkeys * sizeof(int)));
    w.offsets = static_cast<int*>(take(((size_t)g.nkeys + 1) * sizeof(int)));
    w.cursor = static_cast<int*>(take((size_t)g.nkeys * sizeof(int)));
    w.bsum = static_cast<int*>(take((size_t)SCAN_BLOCK * sizeof(int)));
    w.sorted = static_cast<float4*>(take((size_t)(np > 0 ? np : 1) * sizeof(float4)));
    w.tile_mass = static_cast<double*>(take((size_t)g.ntiles * sizeof(double)));
  }
  w.bytes = off;
  return w;
}

int grid_for(int64_t items, int nt, int cap) {
  int64_t g = (items + nt - 1) / nt;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}
}  // namespace

size_t scatter_workspace_bytes(const TileGeom& g, bool direct, int64_t np) {
  return carve(g, direct, np, nullptr).bytes;
}

int scatter_run(const ScatterArgs& a, const TileGeom& g, bool direct, const float* mesh_in, float* mesh_out,
                double* stats, void* ws, size_t ws_bytes, hipStream_t s) {
  ScatterWS w = carve(g, direct, a.np, ws);
  if (ws_bytes < w.bytes) return set_error(PK_E_WORKSPACE, "scatter workspace %zu < %zu bytes", ws_bytes, w.bytes);
  double* st = stats ? stats : w.stats;
  const int dev_cap = 65535;
  const int64_t ncell = (int64_t)a.n * a.n * a.n;
  {
    StageTimer t(ST_FILL, s);
    hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(PK_STATS_LEN), 0, s, st, PK_STATS_LEN, 0.0);
    if (!direct) hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(g.nkeys, 256, 4096)), dim3(256), 0, s, w.counts,
                                    (int64_t)g.nkeys, 0);
    else hipLaunchKernelGGL(k_copy_or_zero, dim3(grid_for(ncell, 256, 16384)), dim3(256), 0, s, mesh_in, mesh_out, ncell);
    int e = check_hip(hipGetLastError(), "scatter fill launch");
    if (e) return e;
  }
  if (direct) {
    if (a.np > 0) {
      StageTimer t(ST_ATOMIC, s);
      hipLaunchKernelGGL(k_atomic_deposit, dim3(grid_for(a.np, DEP_NT, dev_cap)), dim3(DEP_NT), 0, s, a,
                         (const float4*)nullptr, (const int*)nullptr, mesh_out, st);
      int e = check_hip(hipGetLastError(), "atomic deposit launch");
      if (e) return e;
    }
    StageTimer t(ST_MASS, s);
    hipLaunchKernelGGL(k_mass_finalize, dim3(1), dim3(DEP_NT), 0, s, (const double*)nullptr, 0, st, (double)ncell);
    return check_hip(hipGetLastError(), "mass finalize launch");
  }
  const int pgrid = grid_for(a.np, 256, 8192);
  if (a.np > 0) {
    StageTimer t(ST_SORT_COUNT, s);
    hipLaunchKernelGGL(k_sort_count, dim3(pgrid), dim3(256), 0, s, a, g, w.counts);
    int e = check_hip(hipGetLastError(), "sort count launch");
    if (e) return e;
  }
  {
    StageTimer t(ST_SCAN, s);
    const int nblk = (g.nkeys + SCAN_BLOCK - 1) / SCAN_BLOCK;
    hipLaunchKernelGGL(k_scan_local, dim3(nblk), dim3(SCAN_NT), 0, s, w.counts, w.offsets, w.bsum, g.nkeys);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_NT), 0, s, w.bsum, nblk, w.offsets + g.nkeys);
    hipLaunchKernelGGL(k_scan_add, dim3(nblk), dim3(SCAN_NT), 0, s, w.offsets, w.cursor, w.bsum, g.nkeys);
    int e = check_hip(hipGetLastError(), "scan launch");
    if (e) return e;
  }
  if (a.np > 0) {
    StageTimer t(ST_SORT_PLACE, s);
    hipLaunchKernelGGL(k_sort_place, dim3(pgrid), dim3(256), 0, s, a, g, w.cursor, w.sorted);
    int e = check_hip(hipGetLastError(), "sort place launch");
    if (e) return e;
  }
  {
    StageTimer t(ST_DEPOSIT, s);
    const size_t lds = (size_t)g.td[0] * g.td[1] * g.td[2] * sizeof(float);
    hipLaunchKernelGGL(k_deposit_tiles, dim3(g.ntiles), dim3(DEP_NT), lds, s, a, g, w.offsets,
                       (const float4*)w.sorted, mesh_in, mesh_out, w.tile_mass);
    int e = check_hip(hipGetLastError(), "tile deposit launch");
    if (e) return e;
  }
  if (a.stencil == PK_STENCIL_CIC_REF && a.np > 0) {
    StageTimer t(ST_ATOMIC, s);
    hipLaunchKernelGGL(k_atomic_deposit, dim3(64), dim3(DEP_NT), 0, s, a, (const float4*)w.sorted,
                       (const int*)(w.offsets + g.nkeys - 1), mesh_out, st);
    int e = check_hip(hipGetLastError(), "irregular deposit launch");
    if (e) return e;
  }
  StageTimer t(ST_MASS, s);
  hipLaunchKernelGGL(k_mass_finalize, dim3(1), dim3(DEP_NT), 0, s, (const double*)w.tile_mass, g.ntiles, st,
                     (double)ncell);
  return check_hip(hipGetLastError(), "mass finalize launch");
}

}  // namespace pk
