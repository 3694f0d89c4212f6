// Backward of the mass assignment: the gather that transposes the scatter.
//
// Forward (src/mas.py:88-153 and the textbook CIC / TSC stencils):
//   rho(x) = sum_i w_i W(x - p_i),   delta(x) = rho(x) n^3 / M - 1,   M = sum_x rho(x).
// With g(x) = dLoss/d delta(x) (from the spectral adjoint), S = n^3 / M and
// gbar = sum_x g(x) rho(x) / M:
//   dLoss/dp_i = S w_i [ sum_c g(c) dW_c/dp_i - gbar sum_c dW_c/dp_i ]
//   dLoss/dw_i = S     [ sum_c g(c) W_c       - gbar sum_c W_c       ]
// The gbar terms carry the dependence of the mean density on the particles; for the
// mass-conserving stencils sum_c W_c = 1 and its derivative vanishes, for the reference
// stencil (whose :149 corner weight does not conserve mass) it does not.
// Positions in physical units: d pos / d x = inv = 1 / (box / n) in float32, as the forward.
#include "pk_particles.h"

namespace pk {

// Every stencil corner of one particle with its unit-weight W and dW/d(pos): same corners,
// wrap and drop rules as for_each_corner.  f(ix, iy, iz, W, dWx, dWy, dWz).
template <typename F>
__device__ __forceinline__ void for_each_corner_grad(const ScatterArgs& a, float px, float py, float pz, F&& f) {
  const int n = a.n;
  if (a.stencil == PK_STENCIL_CIC_REF) {
    const RefAxis X = ref_axis(px, n, a.wrap_mode), Y = ref_axis(py, n, a.wrap_mode), Z = ref_axis(pz, n, a.wrap_mode);
    // dd = pos - trunc(pos) has slope 1 except where the scan-style no-wrap rule zeroes it
    // (src/mas.py:34-38): there dd = 0 while md keeps 1 - d
    auto slope = [&](float pos) {
      const int i = (int)pos;
      return (a.wrap_mode == PK_WRAP_SCAN_FALSE && i + 1 >= n) ? 0.0f : 1.0f;
    };
    const float sx = slope(px), sy = slope(py), sz = slope(pz);
    // corner weights of src/mas.py:142-151 (the :149 corner takes md_y, not dd_y)
    if (X.vb && Y.vb && Z.vb) f(X.b, Y.b, Z.b, X.md * Y.md * Z.md, -Y.md * Z.md, -X.md * Z.md, -X.md * Y.md);
    if (X.vp && Y.vb && Z.vb) f(X.p, Y.b, Z.b, X.dd * Y.md * Z.md, sx * Y.md * Z.md, -X.dd * Z.md, -X.dd * Y.md);
    if (X.vb && Y.vp && Z.vb) f(X.b, Y.p, Z.b, X.md * Y.dd * Z.md, -Y.dd * Z.md, X.md * sy * Z.md, -X.md * Y.dd);
    if (X.vb && Y.vb && Z.vp) f(X.b, Y.b, Z.p, X.md * Y.md * Z.dd, -Y.md * Z.dd, -X.md * Z.dd, X.md * Y.md * sz);
    if (X.vp && Y.vp && Z.vb) f(X.p, Y.p, Z.b, X.dd * Y.dd * Z.md, sx * Y.dd * Z.md, X.dd * sy * Z.md, -X.dd * Y.dd);
    if (X.vp && Y.vb && Z.vp) f(X.p, Y.b, Z.p, X.dd * Y.md * Z.dd, sx * Y.md * Z.dd, -X.dd * Z.dd, X.dd * Y.md * sz);
    if (X.vb && Y.vp && Z.vp) f(X.b, Y.p, Z.p, X.md * Y.md * Z.dd, -Y.md * Z.dd, -X.md * Z.dd, X.md * Y.md * sz);
    if (X.vp && Y.vp && Z.vp) f(X.p, Y.p, Z.p, X.dd * Y.dd * Z.dd, sx * Y.dd * Z.dd, X.dd * sy * Z.dd, X.dd * Y.dd * sz);
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
      const int bx = s >> 2, by = (s >> 1) & 1, bz = s & 1;
      const float wx = bx ? w1[0] : w0[0], wy = by ? w1[1] : w0[1], wz = bz ? w1[2] : w0[2];
      const float gx = bx ? 1.0f : -1.0f, gy = by ? 1.0f : -1.0f, gz = bz ? 1.0f : -1.0f;
      f(bx ? c1[0] : c0[0], by ? c1[1] : c0[1], bz ? c1[2] : c0[2], wx * wy * wz, gx * wy * wz, wx * gy * wz,
        wx * wy * gz);
    }
  } else {  // TSC: 1/2 (1/2 - d)^2, 3/4 - d^2, 1/2 (1/2 + d)^2 and slopes -(1/2 - d), -2 d, 1/2 + d
    const float pos[3] = {px, py, pz};
    int c[3][3];
    float wt[3][3], dw[3][3];
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
      dw[ax][0] = -hm;
      dw[ax][1] = -2.0f * d;
      dw[ax][2] = hp;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k)
          f(c[0][i], c[1][j], c[2][k], wt[0][i] * wt[1][j] * wt[2][k], dw[0][i] * wt[1][j] * wt[2][k],
            wt[0][i] * dw[1][j] * wt[2][k], wt[0][i] * wt[1][j] * dw[2][k]);
  }
}

constexpr int GAT_NT = 256;
constexpr int GAT_MAX_BLOCKS = 4096;
constexpr int COEF_NT = 256;

__global__ void __launch_bounds__(COEF_NT) k_adjoint_coef(const double* __restrict__ grad_pk,
                                                          const int64_t* __restrict__ counts, int nb, double norm,
                                                          float* __restrict__ coef) {
  const int b = blockIdx.x * COEF_NT + threadIdx.x;
  if (b >= nb) return;
  const double n = (double)counts[b];
#pragma unroll
  for (int l = 0; l < 3; ++l)
    coef[3 * b + l] = n > 0.0 ? (float)(grad_pk[3 * b + l] * norm * (double)(4 * l + 1) / n) : 0.0f;
}

// pass 1: per-block FP64 partials of sum_i w_i sum_c W_c g(c) (= sum_x g rho)
__global__ void __launch_bounds__(GAT_NT) k_gather_grho(ScatterArgs a, const float* __restrict__ g,
                                                        double* __restrict__ part) {
  __shared__ double red[GAT_NT];
  const int n = a.n;
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)GAT_NT + threadIdx.x; i < a.np; i += (int64_t)gridDim.x * GAT_NT) {
    float px, py, pz, w;
    load_particle(a, i, px, py, pz, w);
    if (!pos_ok(px, py, pz)) continue;
    float s = 0.0f;
    for_each_corner_grad(a, px, py, pz, [&](int ix, int iy, int iz, float W, float, float, float) {
      s += W * g[((size_t)ix * n + iy) * n + iz];
    });
    acc += (double)w * (double)s;
  }
  const double t = block_sum_f64(acc, red, GAT_NT);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// gbar = sum(part) / M, fixed order
__global__ void __launch_bounds__(GAT_NT) k_gather_gbar(const double* __restrict__ part, int nparts,
                                                        const double* __restrict__ stats, double* __restrict__ gbar) {
  __shared__ double red[GAT_NT];
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += GAT_NT) s += part[i];
  const double t = block_sum_f64(s, red, GAT_NT);
  if (threadIdx.x == 0) *gbar = stats[0] != 0.0 ? t / stats[0] : 0.0;
}

// pass 2: per-particle gradients
__global__ void __launch_bounds__(GAT_NT) k_gather_grad(ScatterArgs a, const float* __restrict__ g,
                                                        const double* __restrict__ stats,
                                                        const double* __restrict__ gbar_p, float* __restrict__ grad_pos,
                                                        float* __restrict__ grad_w) {
  const int n = a.n;
  const float S = (float)stats[1];
  const float gbar = (float)*gbar_p;
  for (int64_t i = blockIdx.x * (int64_t)GAT_NT + threadIdx.x; i < a.np; i += (int64_t)gridDim.x * GAT_NT) {
    float px, py, pz, w;
    load_particle(a, i, px, py, pz, w);
    float ax = 0.0f, ay = 0.0f, az = 0.0f, bx = 0.0f, by = 0.0f, bz = 0.0f, s0 = 0.0f, sw = 0.0f;
    if (pos_ok(px, py, pz)) {
      for_each_corner_grad(a, px, py, pz, [&](int ix, int iy, int iz, float W, float dx, float dy, float dz) {
        const float gv = g[((size_t)ix * n + iy) * n + iz];
        ax += gv * dx;
        ay += gv * dy;
        az += gv * dz;
        bx += dx;
        by += dy;
        bz += dz;
        s0 += gv * W;
        sw += W;
      });
    }
    const float f = S * w * a.inv;
    grad_pos[3 * i + 0] = f * (ax - gbar * bx);
    grad_pos[3 * i + 1] = f * (ay - gbar * by);
    grad_pos[3 * i + 2] = f * (az - gbar * bz);
    if (grad_w) grad_w[i] = S * (s0 - gbar * sw);
  }
}

namespace {
int gather_grid(int64_t np) {
  int64_t g = (np + GAT_NT - 1) / GAT_NT;
  if (g < 1) g = 1;
  if (g > GAT_MAX_BLOCKS) g = GAT_MAX_BLOCKS;
  return (int)g;
}
}  // namespace

int adjoint_coefficients(const double* grad_pk, const int64_t* counts, int nb, double norm, float* coef,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_adjoint_coef, dim3((nb + COEF_NT - 1) / COEF_NT), dim3(COEF_NT), 0, s, grad_pk, counts, nb,
                     norm, coef);
  return check_hip(hipGetLastError(), "adjoint coefficient launch");
}

size_t gather_workspace_bytes(int64_t np) {
  return align_up((size_t)gather_grid(np) * sizeof(double)) + align_up(sizeof(double));
}

int gather_gradients(const ScatterArgs& a, const float* gmesh, const double* stats, float* grad_pos, float* grad_w,
                     void* ws, size_t ws_bytes, hipStream_t s) {
  if (ws_bytes < gather_workspace_bytes(a.np)) return set_error(PK_E_WORKSPACE, "gather workspace too small");
  const int G = gather_grid(a.np);
  double* part = static_cast<double*>(ws);
  double* gbar = reinterpret_cast<double*>(static_cast<char*>(ws) + align_up((size_t)G * sizeof(double)));
  StageTimer t(ST_GATHER, s);
  hipLaunchKernelGGL(k_gather_grho, dim3(G), dim3(GAT_NT), 0, s, a, gmesh, part);
  hipLaunchKernelGGL(k_gather_gbar, dim3(1), dim3(GAT_NT), 0, s, (const double*)part, G, stats, gbar);
  hipLaunchKernelGGL(k_gather_grad, dim3(G), dim3(GAT_NT), 0, s, a, gmesh, stats, (const double*)gbar, grad_pos,
                     grad_w);
  return check_hip(hipGetLastError(), "gather launch");
}

}  // namespace pk
