// Device-side particle helpers shared by the assignment kernels: float32 position
// arithmetic of the reference, stencil corners and the row-bundle tile keys.
//
// Reference arithmetic (src/mas.py:100-113, 138-140), float32 on float32 inputs:
//   pos = (x - xmin) * inv, inv = 1 / (box / n);  ix = int32(pos) (truncation);
//   dd = pos - ix;  md = 1 - dd.
#pragma once
#include "pk_common.h"
#include "pk_internal.h"

namespace pk {

constexpr float POS_LIMIT = 1073741824.0f;  // |pos| >= 2^30 or non-finite: particle dropped
constexpr int MAX_TILE_BUCKETS = 16;          // buckets a row-bundle tile reads (incl. irregular)

__device__ __forceinline__ int floor_mod(int a, int n) {
  // indexes of positions inside the box (or one period above) take two instructions; the
  // integer division only runs for lanes that need it
  if ((unsigned)a < (unsigned)(2 * n)) return a >= n ? a - n : a;
  int r = a % n;
  return r < 0 ? r + n : r;
}

// AOS / W: the input layout (x, y, z interleaved) and the presence of weights, fixed at
// compile time (0 / 1) or read from a (-1).  Fixed values leave no branch between the loads
// of a thread's particles, so they are all in flight together; the position arithmetic
// follows the loads.  STREAM: non-temporal loads, for the two reads of the positions in an
// estimate (count, coarse place), so that they do not displace in the caches what the next
// kernels read (keeping them cached between the two reads measured slower overall).
template <int AOS = -1, int W = -1, bool STREAM = false>
__device__ __forceinline__ void load_particle(const ScatterArgs& a, int64_t i, float& px, float& py, float& pz,
                                              float& w) {
  const bool aos = AOS >= 0 ? AOS == 1 : a.aos != 0;
  const bool hw = W >= 0 ? W == 1 : a.w != nullptr;
  auto ld = [](const float* p) -> float {
    if constexpr (STREAM) return __builtin_nontemporal_load(p);
    else return *p;
  };
  float x, y, z;
  if (aos) {   // one 12-byte load per particle
    typedef float f3 __attribute__((ext_vector_type(3)));
    typedef float f3a __attribute__((ext_vector_type(3), aligned(4)));
    f3 v;
    if constexpr (STREAM) v = __builtin_nontemporal_load(reinterpret_cast<const f3a*>(a.x + 3 * i));
    else __builtin_memcpy(&v, a.x + 3 * i, sizeof(float) * 3);
    x = v.x;
    y = v.y;
    z = v.z;
  } else {
    x = ld(a.x + i * a.sx);
    y = ld(a.y + i * a.sy);
    z = ld(a.z + i * a.sz);
  }
  w = hw ? ld(a.w + i * a.sw) : 1.0f;
  px = (x - a.xmin) * a.inv;
  py = (y - a.ymin) * a.inv;
  pz = (z - a.zmin) * a.inv;
}

__device__ __forceinline__ bool pos_ok(float px, float py, float pz) {
  return fabsf(px) < POS_LIMIT && fabsf(py) < POS_LIMIT && fabsf(pz) < POS_LIMIT;  // false for NaN
}

// one axis of the reference stencil: base index b, +1 index p (wrap rule of `mode`),
// validity after JAX index normalisation (negative -> +n, outside [0, n) -> dropped)
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
    p = (p >= n) ? 0 : p;                  // src/mas.py:125 (the ddx test at :126 never fires)
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

// home cell per axis: base cell (reference stencil, CIC) or nearest grid point (TSC)
__device__ __forceinline__ int home_cell(float pos, int n, int stencil) {
  if (stencil == PK_STENCIL_TSC) return floor_mod((int)floorf(pos + 0.5f), n);
  if (stencil == PK_STENCIL_CIC) return floor_mod((int)floorf(pos), n);
  const int i = (int)pos;
  return i < 0 ? i + n : i;
}

// Visit every stencil corner of one particle: f(ix, iy, iz, weight), indexes in [0, n).
// The reference stencil applies the JAX drop rule per corner.
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
    // weights ((wx w) wy) wz: the plane and row factors are shared by their corners
#pragma unroll
    for (int sx = 0; sx < 2; ++sx) {
      const float ax = (sx ? w1[0] : w0[0]) * w;
#pragma unroll
      for (int sy = 0; sy < 2; ++sy) {
        const float axy = ax * (sy ? w1[1] : w0[1]);
#pragma unroll
        for (int sz = 0; sz < 2; ++sz)
          f(sx ? c1[0] : c0[0], sy ? c1[1] : c0[1], sz ? c1[2] : c0[2], axy * (sz ? w1[2] : w0[2]));
      }
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
    // weights ((wx w) wy) wz, as for CIC
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float ax = wt[0][i] * w;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const float axy = ax * wt[1][j];
#pragma unroll
        for (int k = 0; k < 3; ++k) f(c[0][i], c[1][j], c[2][k], axy * wt[2][k]);
      }
    }
  }
}

// The corners of one particle that lie on the x plane `xt`: f(iy, iz, weight).  Same
// weights and drop rules as for_each_corner, restricted to one plane, so a tile that owns
// a single x layer issues 4 (CIC variants) or 9 (TSC) deposits per particle, not 8 or 27.
// ST >= 0 fixes the stencil at compile time (the run-time a.stencil is then ignored).
template <int ST = -1, typename F>
__device__ __forceinline__ void for_each_corner_on_x(const ScatterArgs& a, float px, float py, float pz, float w,
                                                     int xt, F&& f) {
  const int n = a.n;
  const int stencil = ST >= 0 ? ST : a.stencil;
  if (stencil == PK_STENCIL_CIC_REF) {
    const RefAxis X = ref_axis(px, n, a.wrap_mode), Y = ref_axis(py, n, a.wrap_mode), Z = ref_axis(pz, n, a.wrap_mode);
    if (X.vb && X.b == xt) {   // src/mas.py:142, :144, :145, :149 (the :149 weight keeps mdy)
      if (Y.vb && Z.vb) f(Y.b, Z.b, X.md * Y.md * Z.md * w);
      if (Y.vp && Z.vb) f(Y.p, Z.b, X.md * Y.dd * Z.md * w);
      if (Y.vb && Z.vp) f(Y.b, Z.p, X.md * Y.md * Z.dd * w);
      if (Y.vp && Z.vp) f(Y.p, Z.p, X.md * Y.md * Z.dd * w);
    }
    if (X.vp && X.p == xt) {   // src/mas.py:143, :147, :148, :151
      if (Y.vb && Z.vb) f(Y.b, Z.b, X.dd * Y.md * Z.md * w);
      if (Y.vp && Z.vb) f(Y.p, Z.b, X.dd * Y.dd * Z.md * w);
      if (Y.vb && Z.vp) f(Y.b, Z.p, X.dd * Y.md * Z.dd * w);
      if (Y.vp && Z.vp) f(Y.p, Z.p, X.dd * Y.dd * Z.dd * w);
    }
  } else if (stencil == PK_STENCIL_CIC) {
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
    for (int sx = 0; sx < 2; ++sx) {
      if ((sx ? c1[0] : c0[0]) != xt) continue;
      const float ax = (sx ? w1[0] : w0[0]) * w;   // same products as for_each_corner
#pragma unroll
      for (int sy = 0; sy < 2; ++sy) {
        const float axy = ax * (sy ? w1[1] : w0[1]);
#pragma unroll
        for (int sz = 0; sz < 2; ++sz) f(sy ? c1[1] : c0[1], sz ? c1[2] : c0[2], axy * (sz ? w1[2] : w0[2]));
      }
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
    for (int i = 0; i < 3; ++i) {
      if (c[0][i] != xt) continue;
      const float ax = wt[0][i] * w;   // same products as for_each_corner
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const float axy = ax * wt[1][j];
#pragma unroll
        for (int k = 0; k < 3; ++k) f(c[1][j], c[2][k], axy * wt[2][k]);
      }
    }
  }
}

// The corners on plane xt of a reference-stencil particle whose truncated base index is in
// [0, n) on every axis (every bucket but the irregular one).  Then every index is valid and
// the four wrap rules give the same +1 neighbour: 0 at the upper edge, else i + 1; only
// cic_mas with wrap=False also zeroes the fraction at the upper edge (src/mas.py:34).  Same
// weights and products, left to right, as for_each_corner_on_x (incl. the :149 corner),
// without its per-corner validity tests.
template <typename F>
__device__ __forceinline__ void ref_corners_on_x_regular(const ScatterArgs& a, float px, float py, float pz,
                                                         float w, int xt, F&& f) {
  const int n = a.n;
  const bool scan_false = a.wrap_mode == PK_WRAP_SCAN_FALSE;
  auto axis = [&](float pos, int& b, int& p, float& md, float& dd) {
    const int i = (int)pos;   // truncation toward zero; 0 <= i < n here
    const float d = pos - (float)i;
    md = 1.0f - d;
    const bool top = i + 1 >= n;
    dd = (scan_false && top) ? 0.0f : d;
    b = i;
    p = top ? 0 : i + 1;
  };
  int xb, xp, yb, yp, zb, zp;
  float xm, xd, ym, yd, zm, zd;
  axis(px, xb, xp, xm, xd);
  axis(py, yb, yp, ym, yd);
  axis(pz, zb, zp, zm, zd);
  if (xb == xt) {   // src/mas.py:142, :144, :145, :149
    f(yb, zb, xm * ym * zm * w);
    f(yp, zb, xm * yd * zm * w);
    f(yb, zp, xm * ym * zd * w);
    f(yp, zp, xm * ym * zd * w);
  }
  if (xp == xt) {   // src/mas.py:143, :147, :148, :151
    f(yb, zb, xd * ym * zm * w);
    f(yp, zb, xd * yd * zm * w);
    f(yb, zp, xd * ym * zd * w);
    f(yp, zp, xd * yd * zd * w);
  }
}

// Sorted particle records: float4(xpos, ypos, zpos, w) when weights are given, else
// float3(xpos, ypos, zpos) packed at 12 bytes (w = 1).  Record arrays hold 16 bytes per
// particle, so a 16-byte read of the last packed record stays inside the array.
__device__ __forceinline__ void store_record(void* sorted, size_t i, bool has_w, float px, float py, float pz,
                                             float w) {
  typedef float f3 __attribute__((ext_vector_type(3), aligned(4)));
  float* p = static_cast<float*>(sorted) + i * (has_w ? 4 : 3);
  const f3 v = {px, py, pz};
  *reinterpret_cast<f3*>(p) = v;
  if (has_w) p[3] = w;
}

// one 16-byte load at the record's stride for both layouts: no branch between the loads of
// several records.  The fourth word of a packed record belongs to the next one: callers
// that defer the weight select (record_weight) past their other loads keep every load in
// flight.
__device__ __forceinline__ float4 load_record_raw(const void* sorted, size_t i, bool has_w) {
  typedef float f4 __attribute__((ext_vector_type(4), aligned(4)));
  const float* p = static_cast<const float*>(sorted) + i * (has_w ? 4 : 3);
  const f4 v = *reinterpret_cast<const f4*>(p);
  return make_float4(v.x, v.y, v.z, v.w);
}
// The same as a non-temporal load, for the last read of a record array in an estimate (the
// staged records in the fine place, the sorted ones in the fused deposit): the 0.24 GB of
// records then do not displace in the caches what the next kernels read.
__device__ __forceinline__ float4 load_record_stream(const void* sorted, size_t i, bool has_w) {
  typedef float f4 __attribute__((ext_vector_type(4), aligned(4)));
  const float* p = static_cast<const float*>(sorted) + i * (has_w ? 4 : 3);
  const f4 v = __builtin_nontemporal_load(reinterpret_cast<const f4*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float record_weight(const float4& r, bool has_w) { return has_w ? r.w : 1.0f; }
__device__ __forceinline__ float4 load_record(const void* sorted, size_t i, bool has_w) {
  float4 r = load_record_raw(sorted, i, has_w);
  r.w = record_weight(r, has_w);
  return r;
}

// Row-bundle tiles: tile (xt, yt) owns cells x = xt, y in [yt*ty, yt*ty + ty), all z.
// Key = tile * ncls + y-class, where the y-class marks the tile rows whose stencil spills
// into the neighbouring tile (CIC variants: 1 = last row; TSC: 1 = first, 2 = last row).
// The x neighbours are read whole (one x-layer per tile).  Last key: irregular particles
// (reference stencil with an out-of-range base index); -1: dropped.
// plane (optional): the key's x plane hx, or n for the irregular key (keys are x-plane major).
// ST >= 0 fixes the stencil at compile time (the run-time a.stencil is then ignored).
template <int ST = -1>
__device__ __forceinline__ int row_tile_key(const ScatterArgs& a, const TileGeom& g, float px, float py, float pz,
                                            int* plane = nullptr) {
  if (!pos_ok(px, py, pz)) return -1;
  const int n = a.n;
  const int stencil = ST >= 0 ? ST : a.stencil;
  if (stencil == PK_STENCIL_CIC_REF) {
    // irregular: a truncated base index outside [0, n) on any axis (positions below -1 cell
    // or at or above n), whose corners need the per-corner JAX index rules
    const int ix = (int)px, iy = (int)py, iz = (int)pz;
    if ((unsigned)ix >= (unsigned)n || (unsigned)iy >= (unsigned)n || (unsigned)iz >= (unsigned)n) {
      if (plane) *plane = n;
      return g.nkeys - 1;
    }
  }
  const int hx = home_cell(px, n, stencil);
  const int hy = home_cell(py, n, stencil);
  const int yt = (int)__umulhi((unsigned)hy, g.ty_mul);   // hy / ty
  const int ly = hy - yt * g.ty;
  int cls;
  if (stencil == PK_STENCIL_TSC) cls = (ly == 0) ? 1 : ((ly == g.ty - 1) ? 2 : 0);
  else cls = (ly == g.ty - 1) ? 1 : 0;
  if (plane) *plane = hx;
  return (hx * g.nyt + yt) * g.ncls + cls;
}

// The x plane of row_tile_key's key (n for the irregular key), or -1 where it returns -1:
// all the coarse placing pass needs, without the key's y tile and class.
template <int ST = -1>
__device__ __forceinline__ int row_plane(const ScatterArgs& a, float px, float py, float pz) {
  if (!pos_ok(px, py, pz)) return -1;
  const int n = a.n;
  const int stencil = ST >= 0 ? ST : a.stencil;
  if (stencil == PK_STENCIL_CIC_REF) {
    const int ix = (int)px, iy = (int)py, iz = (int)pz;
    if ((unsigned)ix >= (unsigned)n || (unsigned)iy >= (unsigned)n || (unsigned)iz >= (unsigned)n) return n;
  }
  return home_cell(px, n, stencil);
}

// Buckets whose particles can deposit into tile (xt, yt): filled into key[] (<= 16).
__host__ __device__ __forceinline__ int row_tile_buckets(const TileGeom& g, int n, bool tsc, int xt, int yt, int* keys) {
  int xs[3], nxs = 0;
  xs[nxs++] = xt;
  if (n > 1) xs[nxs++] = (xt + n - 1) % n;                 // lower x neighbour (all of it)
  if (tsc && n > 2) xs[nxs++] = (xt + 1) % n;              // upper x neighbour (TSC)
  // y: (tile, class) pairs
  int yc[8], ycl[8], ny = 0;
  for (int c = 0; c < g.ncls; ++c) { yc[ny] = yt; ycl[ny] = c; ++ny; }
  if (g.nyt >= 2) {
    const int lo = (yt + g.nyt - 1) % g.nyt, hi = (yt + 1) % g.nyt;
    if (tsc) {
      if (g.nyt >= 3) {
        yc[ny] = lo; ycl[ny] = 2; ++ny;   // lower tile, last row
        yc[ny] = hi; ycl[ny] = 1; ++ny;   // upper tile, first row
      } else {
        yc[ny] = lo; ycl[ny] = 1; ++ny;   // the other tile's first and last rows
        yc[ny] = lo; ycl[ny] = 2; ++ny;
      }
    } else {
      yc[ny] = lo; ycl[ny] = 1; ++ny;
    }
  }
  int nk = 0;
  for (int i = 0; i < nxs; ++i)
    for (int j = 0; j < ny; ++j) keys[nk++] = (xs[i] * g.nyt + yc[j]) * g.ncls + ycl[j];
  return nk;
}

// Number of buckets row_tile_buckets lists (the same for every tile), plus the irregular one.
__host__ __device__ __forceinline__ int tile_bucket_count(const TileGeom& g, int n, bool tsc, bool irregular) {
  const int nxs = 1 + (n > 1 ? 1 : 0) + (tsc && n > 2 ? 1 : 0);
  const int ny = g.ncls + (g.nyt >= 2 ? (tsc ? 2 : 1) : 0);
  return nxs * ny + (irregular ? 1 : 0);
}

// Key q of tile (xt, yt) in row_tile_buckets' order; q = nxs * ny is the irregular bucket.
// Each lane computes its own key, so no lane waits for another before loading key starts.
__host__ __device__ __forceinline__ int tile_bucket_key(const TileGeom& g, int n, bool tsc, int xt, int yt, int q) {
  const int nxs = 1 + (n > 1 ? 1 : 0) + (tsc && n > 2 ? 1 : 0);
  const int ny = g.ncls + (g.nyt >= 2 ? (tsc ? 2 : 1) : 0);
  if (q >= nxs * ny) return g.nkeys - 1;
  const int i = q / ny, jj = q - (q / ny) * ny;
  const int x = i == 0 ? xt : (i == 1 ? (xt + n - 1) % n : (xt + 1) % n);
  int yc = yt, ycl = jj;
  if (jj >= g.ncls) {
    const int e = jj - g.ncls;
    const int lo = (yt + g.nyt - 1) % g.nyt, hi = (yt + 1) % g.nyt;
    if (tsc && g.nyt >= 3) {
      yc = e == 0 ? lo : hi;
      ycl = e == 0 ? 2 : 1;
    } else if (tsc) {
      yc = lo;
      ycl = e == 0 ? 1 : 2;
    } else {
      yc = lo;
      ycl = 1;
    }
  }
  return (x * g.nyt + yc) * g.ncls + ycl;
}

// Block-cooperative setup of the buckets of tile (xt, yt) (plus the irregular bucket when
// `irregular`): keys, first record and exclusive prefix of sizes in LDS.  Ends with a barrier.
__device__ __forceinline__ void gather_buckets(const TileGeom& g, int n, bool tsc, bool irregular, int xt, int yt,
                                               const int* __restrict__ key_start, int* s_key, int* s_base,
                                               int* s_pre, int* s_nbk) {
  const int tid = threadIdx.x;
  if (tid == 0) {
    int nk = row_tile_buckets(g, n, tsc, xt, yt, s_key);
    if (irregular) s_key[nk++] = g.nkeys - 1;
    *s_nbk = nk;
  }
  __syncthreads();
  const int nk = *s_nbk;
  if (tid < nk) {
    const int lo = key_start[s_key[tid]];
    s_base[tid] = lo;
    s_pre[tid + 1] = key_start[s_key[tid] + 1] - lo;   // size, prefixed below
  }
  __syncthreads();
  if (tid == 0) {
    s_pre[0] = 0;
    for (int b = 1; b <= nk; ++b) s_pre[b] += s_pre[b - 1];
  }
  __syncthreads();
}

// 64-bit fixed-point deposit value (two's complement) of a weight already multiplied by
// the scale S = 2^(30 - e) with 2^e > max|w|: every corner weight is <= 1, so |v| < 2^30,
// rint(v) fits an int32 and the conversion is one round and one convert (a float -> int64
// conversion would take ten instructions per corner).
__device__ __forceinline__ unsigned long long to_fixed(float v) {
  return (unsigned long long)(long long)(int)__builtin_rintf(v);
}
__device__ __forceinline__ unsigned long long to_fixed(float wt, float scale) { return to_fixed(wt * scale); }

// Deposit every record of the tile's buckets (set up by gather_buckets) whose corners lie
// on plane xt and rows [y0, y0 + ty): f(ly, iz, weight).  The bucket of a flat index is
// found with a register copy of the (<= 17) prefix values.
template <typename F>
__device__ __forceinline__ void deposit_buckets(const ScatterArgs& a, const TileGeom& g, const void* sorted,
                                                const int* s_base, const int* s_pre, int nbk, int nt, int xt, int y0,
                                                F&& f) {
  int pre[MAX_TILE_BUCKETS + 1];
#pragma unroll
  for (int b = 0; b <= MAX_TILE_BUCKETS; ++b) pre[b] = b <= nbk ? s_pre[b] : 0x7fffffff;
  const int total = pre[nbk < MAX_TILE_BUCKETS ? nbk : MAX_TILE_BUCKETS];
  const bool has_w = a.w != nullptr;
  constexpr int U = 4;   // independent records per thread and iteration (loads in flight)
  for (int idx0 = threadIdx.x; idx0 < total; idx0 += nt * U) {
    float4 r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int idx = idx0 + u * nt;
      if (idx < total) {
        int b = 0;
#pragma unroll
        for (int j = 1; j < MAX_TILE_BUCKETS; ++j) b += (pre[j] <= idx) ? 1 : 0;
        r[u] = load_record(sorted, (size_t)s_base[b] + (idx - pre[b]), has_w);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (idx0 + u * nt >= total) continue;
      auto add = [&](int iy, int iz, float wt) {
        const int ly = iy - y0;
        if ((unsigned)ly < (unsigned)g.ty) f(ly, iz, wt);
      };
      // the tile buckets hold no irregular record (those are added with global atomics)
      if (a.stencil == PK_STENCIL_CIC_REF) ref_corners_on_x_regular(a, r[u].x, r[u].y, r[u].z, r[u].w, xt, add);
      else for_each_corner_on_x(a, r[u].x, r[u].y, r[u].z, r[u].w, xt, add);
    }
  }
}

}  // namespace pk
