// extern "C" entry points of libpowspec_amd.so (declared in include/pk_api.h): plans,
// workspace carving, argument validation, error state and per-stage event timing.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "pk_common.h"
#include "pk_internal.h"

// ------------------------------------------------------------------------------------
// error state (per host thread)
// ------------------------------------------------------------------------------------
namespace {
thread_local char g_err[512] = "";
}

namespace pk {

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return PK_OK;
  return set_error(PK_E_LAUNCH, "%s: %s", what, hipGetErrorString(e));
}

int num_cus(int device) {
  static int cache[64] = {0};
  if (device >= 0 && device < 64 && cache[device]) return cache[device];
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || v < 1) v = 1;
  if (device >= 0 && device < 64) cache[device] = v;
  return v;
}

// ------------------------------------------------------------------------------------
// per-stage timing
// ------------------------------------------------------------------------------------
namespace {
struct Rec {
  int stage;
  hipEvent_t a, b;
};
std::mutex g_tmu;
bool g_timing = false;
std::vector<Rec> g_recs;
thread_local hipEvent_t g_open[PK_NUM_STAGES] = {};
const char* g_stage_names[PK_NUM_STAGES] = {"sort_count", "scan",   "sort_place",  "deposit",
                                            "atomic",     "mass",   "zpass",       "ypass",
                                            "xpass_bin",  "finalize", "xpass_plain", "fill"};
}  // namespace

void timing_begin(int stage, hipStream_t s) {
  if (!g_timing) return;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  (void)hipEventRecord(e, s);
  g_open[stage] = e;
}

void timing_end(int stage, hipStream_t s) {
  if (!g_timing || !g_open[stage]) return;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  (void)hipEventRecord(e, s);
  std::lock_guard<std::mutex> lk(g_tmu);
  g_recs.push_back(Rec{stage, g_open[stage], e});
  g_open[stage] = nullptr;
}

}  // namespace pk

using namespace pk;

// ------------------------------------------------------------------------------------
// plans
// ------------------------------------------------------------------------------------
struct pk_plan {
  int n;
  int device;
  float2* tw;      // e^{-2 pi i q / n}, q < n
  float* wtab;     // [4][n]: (x / sin x)^(2p), x = pi k / n, p = 0..3
};

namespace {
int host_plan_tables(int n, std::vector<float2>& tw, std::vector<float>& wt) {
  tw.resize(n);
  for (int q = 0; q < n; ++q) {
    const double a = -2.0 * M_PI * (double)q / (double)n;
    tw[q] = make_float2((float)cos(a), (float)sin(a));
  }
  wt.resize(4 * (size_t)n);
  for (int p = 0; p < 4; ++p) {
    for (int i = 0; i < n; ++i) {
      const int k = i <= n / 2 ? i : i - n;
      double f = 1.0;
      if (k != 0) {
        const double x = M_PI * (double)k / (double)n;
        f = pow(x / sin(x), 2.0 * p);
      }
      wt[(size_t)p * n + i] = (float)f;
    }
  }
  return PK_OK;
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

int check_n(int n) {
  if (!fft_supported(n)) return set_error(PK_E_INVALID, "unsupported mesh side n=%d (need 8..1024, 2^a or 3*2^a)", n);
  return PK_OK;
}

int check_plan(pk_plan* p) {
  if (!p) return set_error(PK_E_INVALID, "null plan");
  int dev = -1;
  (void)hipGetDevice(&dev);
  if (dev != p->device) return set_error(PK_E_INVALID, "plan created on device %d used on device %d", p->device, dev);
  return PK_OK;
}

int make_args(ScatterArgs* a, const float* x, int64_t sx, const float* y, int64_t sy, const float* z, int64_t sz,
              const float* w, int64_t sw, int64_t np, float xmin, float ymin, float zmin, double box, int n,
              int stencil, int wrap_mode) {
  if (np < 0 || np > 0x7fffffffLL) return set_error(PK_E_INVALID, "particle count %lld out of range", (long long)np);
  if (np > 0 && (!x || !y || !z)) return set_error(PK_E_INVALID, "null position pointer");
  if (sx < 1 || sy < 1 || sz < 1 || (w && sw < 1)) return set_error(PK_E_INVALID, "element strides must be >= 1");
  if (stencil < PK_STENCIL_CIC_REF || stencil > PK_STENCIL_TSC) return set_error(PK_E_INVALID, "bad stencil %d", stencil);
  if (wrap_mode < PK_WRAP_VEC_TRUE || wrap_mode > PK_WRAP_SCAN_FALSE) return set_error(PK_E_INVALID, "bad wrap mode %d", wrap_mode);
  if (n < 2 || n > 4096) return set_error(PK_E_INVALID, "mesh side n=%d out of range", n);
  if (!(box > 0.0)) return set_error(PK_E_INVALID, "box_size must be > 0");
  a->x = x; a->y = y; a->z = z; a->w = w;
  a->sx = sx; a->sy = sy; a->sz = sz; a->sw = sw;
  a->np = np;
  a->xmin = xmin; a->ymin = ymin; a->zmin = zmin;
  // float32 as the reference: bin_size = box / n_bins; inv = 1 / bin_size (src/mas.py:100-101)
  const float bin = (float)box / (float)n;
  a->inv = 1.0f / bin;
  a->n = n;
  a->stencil = stencil;
  a->wrap_mode = wrap_mode;
  return PK_OK;
}

// workspace of pk_powspec_mesh: bin partials + z-pass DC partials (spectrum is an argument)
size_t powspec_ws_bytes(int n, int nb) {
  const int G = bin_grid(n, nb);
  return align_up((size_t)G * (size_t)(nb > 0 ? nb : 1) * 5 * sizeof(double)) +
         align_up((size_t)zpass_grid(n) * sizeof(double));
}

int check_bins(const int32_t* thr, const int32_t* jstart, int nb, const float* pk) {
  if (nb < 1) return set_error(PK_E_INVALID, "need at least one bin (nb=%d)", nb);
  if (nb > (1 << 20)) return set_error(PK_E_INVALID, "too many bins (nb=%d)", nb);
  if (!thr || !jstart || !pk) return set_error(PK_E_INVALID, "null threshold / jstart / pk pointer");
  return PK_OK;
}

// y pass, x pass + binning and finalize on a spectrum whose z pass is done
int finish_powspec(pk_plan* p, float2* spec, const int32_t* thr, const int32_t* jstart, int nb, int window_p,
                   int flags, double box, double shot, const double* dc_part, const double* scale_stats,
                   double* partials, double* sums, int64_t* counts, float* pk, float* nmodes, double* kmean,
                   hipStream_t s) {
  const int n = p->n;
  int e = fft_ypass(n, spec, p->tw, s);
  if (e) return e;
  const int G = bin_grid(n, nb);
  const int jmax = (int)floor(sqrt(3.0 * (n / 2) * (n / 2))) + 1;
  e = fft_xpass_bin(n, spec, p->tw, p->wtab + (size_t)window_p * n, thr, jstart, nb, jmax, flags, partials, G, s);
  if (e) return e;
  const double h = box / ((double)n * (double)n);
  const double norm = h * h * h;  // (box / n^2)^3, src/correlations.py:50-52
  return bin_finalize(partials, G, nb, thr, dc_part, zpass_grid(n), scale_stats, norm, shot, sums, counts, pk,
                      nmodes, (flags & PK_FLAG_KMEAN) ? kmean : nullptr, s);
}

double* dc_partials(void* ws, int n, int nb) {
  return reinterpret_cast<double*>(static_cast<char*>(ws) +
                                   align_up((size_t)bin_grid(n, nb) * (size_t)nb * 5 * sizeof(double)));
}

int run_powspec(pk_plan* p, const float* mesh, const double* stats, const int32_t* thr, const int32_t* jstart,
                int nb, int window_p, int flags, double box, double shot, float2* spec, double* sums,
                int64_t* counts, float* pk, float* nmodes, double* kmean, void* ws, size_t ws_bytes,
                hipStream_t s) {
  const int n = p->n;
  const size_t need = powspec_ws_bytes(n, nb);
  if (ws_bytes < need) return set_error(PK_E_WORKSPACE, "powspec workspace %zu < %zu bytes", ws_bytes, need);
  if (window_p < 0 || window_p > 3) return set_error(PK_E_INVALID, "window exponent p=%d not in 0..3", window_p);
  double* partials = static_cast<double*>(ws);
  double* dc_part = dc_partials(ws, n, nb);
  // with the contrast applied (ZERO_DC) the k = 0 mode is exactly 0 and no DC sum is needed
  const bool keep_dc = (flags & PK_FLAG_ZERO_DC) == 0;
  int e = fft_zpass(n, mesh, spec, p->tw, stats, keep_dc ? dc_part : nullptr, s);
  if (e) return e;
  return finish_powspec(p, spec, thr, jstart, nb, window_p, flags, box, shot, keep_dc ? dc_part : nullptr, nullptr,
                        partials, sums, counts, pk, nmodes, kmean, s);
}
}  // namespace

extern "C" {

int pk_version(void) { return PK_API_VERSION; }

const char* pk_last_error(void) { return g_err; }

int pk_supported_n(int n) { return fft_supported(n) ? 1 : 0; }

int pk_plan_create(int n, pk_plan** out) {
  if (!out) return set_error(PK_E_INVALID, "null plan output pointer");
  *out = nullptr;
  int e = check_n(n);
  if (e) return e;
  std::vector<float2> tw;
  std::vector<float> wt;
  host_plan_tables(n, tw, wt);
  pk_plan* p = new pk_plan();
  p->n = n;
  (void)hipGetDevice(&p->device);
  if (hipMalloc(&p->tw, tw.size() * sizeof(float2)) != hipSuccess) {
    delete p;
    return set_error(PK_E_OOM, "plan twiddle allocation failed");
  }
  if (hipMalloc(&p->wtab, wt.size() * sizeof(float)) != hipSuccess) {
    (void)hipFree(p->tw);
    delete p;
    return set_error(PK_E_OOM, "plan window-table allocation failed");
  }
  e = check_hip(hipMemcpy(p->tw, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice), "plan upload");
  if (!e) e = check_hip(hipMemcpy(p->wtab, wt.data(), wt.size() * sizeof(float), hipMemcpyHostToDevice), "plan upload");
  if (e) {
    (void)hipFree(p->tw);
    (void)hipFree(p->wtab);
    delete p;
    return e;
  }
  *out = p;
  return PK_OK;
}

int pk_plan_destroy(pk_plan* p) {
  if (!p) return PK_OK;
  (void)hipFree(p->tw);
  (void)hipFree(p->wtab);
  delete p;
  return PK_OK;
}

int pk_scatter_workspace_size(int n, int64_t np, int stencil, const int* tile, size_t* bytes) {
  if (!bytes) return set_error(PK_E_INVALID, "null bytes pointer");
  if (n < 2 || n > 4096) return set_error(PK_E_INVALID, "mesh side n=%d out of range", n);
  if (np < 0 || np > 0x7fffffffLL) return set_error(PK_E_INVALID, "particle count out of range");
  TileGeom g{};
  bool direct = false;
  int e = make_tile_geom(n, np, stencil, tile, &g, &direct);
  if (e) return e;
  *bytes = scatter_workspace_bytes(g, direct, np);
  return PK_OK;
}

int pk_scatter(const float* x, int64_t sx, const float* y, int64_t sy, const float* z, int64_t sz, const float* w,
               int64_t sw, int64_t np, float xmin, float ymin, float zmin, double box_size, int n, int stencil,
               int wrap_mode, const int* tile, const float* mesh_in, float* mesh_out, double* stats,
               void* workspace, size_t workspace_bytes, void* stream) {
  ScatterArgs a{};
  int e = make_args(&a, x, sx, y, sy, z, sz, w, sw, np, xmin, ymin, zmin, box_size, n, stencil, wrap_mode);
  if (e) return e;
  if (!mesh_out) return set_error(PK_E_INVALID, "null mesh_out");
  TileGeom g{};
  bool direct = false;
  e = make_tile_geom(n, np, stencil, tile, &g, &direct);
  if (e) return e;
  return scatter_run(a, g, direct, mesh_in, mesh_out, stats, workspace, workspace_bytes, as_stream(stream));
}

int pk_r2c(pk_plan* p, const float* mesh, const double* stats, void* spectrum, void* stream) {
  int e = check_plan(p);
  if (e) return e;
  if (!mesh || !spectrum) return set_error(PK_E_INVALID, "null mesh / spectrum");
  hipStream_t s = as_stream(stream);
  float2* spec = static_cast<float2*>(spectrum);
  e = fft_zpass(p->n, mesh, spec, p->tw, stats, nullptr, s);
  if (!e) e = fft_ypass(p->n, spec, p->tw, s);
  if (!e) e = fft_xpass(p->n, spec, p->tw, s);
  return e;
}

int pk_powspec_workspace_size(pk_plan* p, int nb, size_t* bytes) {
  int e = check_plan(p);
  if (e) return e;
  if (!bytes) return set_error(PK_E_INVALID, "null bytes pointer");
  if (nb < 1) return set_error(PK_E_INVALID, "need at least one bin (nb=%d)", nb);
  *bytes = powspec_ws_bytes(p->n, nb);
  return PK_OK;
}

int pk_powspec_mesh(pk_plan* p, const float* mesh, const double* stats, const int32_t* thr, const int32_t* jstart,
                    int nb, int window_p, int flags, double box_size, double shot_noise, void* spectrum,
                    double* sums, int64_t* counts, float* pk, float* nmodes, double* kmean, void* workspace,
                    size_t workspace_bytes, void* stream) {
  int e = check_plan(p);
  if (e) return e;
  e = check_bins(thr, jstart, nb, pk);
  if (e) return e;
  if (!mesh || !spectrum) return set_error(PK_E_INVALID, "null mesh / spectrum");
  if (!(box_size > 0.0)) return set_error(PK_E_INVALID, "box_size must be > 0");
  return run_powspec(p, mesh, stats, thr, jstart, nb, window_p, flags, box_size, shot_noise,
                     static_cast<float2*>(spectrum), sums, counts, pk, nmodes, kmean, workspace, workspace_bytes,
                     as_stream(stream));
}

namespace {
struct EstimateWS {
  bool fused;
  float* mesh;
  float2* spec;
  double* stats;
  void* scatter_ws;
  size_t scatter_bytes;
  void* powspec_ws;
  size_t powspec_bytes;
  size_t bytes;
};

int carve_estimate(pk_plan* p, int64_t np, int stencil, const int* tile, int nb, void* base, EstimateWS* w,
                   TileGeom* g, bool* direct) {
  const int n = p->n;
  int e = make_tile_geom(n, np, stencil, tile, g, direct);
  if (e) return e;
  w->fused = !*direct && fused_supported(n, *g);
  char* b = static_cast<char*>(base);
  size_t off = 0;
  auto take = [&](size_t bytes) -> void* {
    void* q = b ? b + off : nullptr;
    off += align_up(bytes);
    return q;
  };
  const size_t n3 = (size_t)n * n * n;
  w->mesh = w->fused ? nullptr : static_cast<float*>(take(n3 * sizeof(float)));
  w->spec = static_cast<float2*>(take((size_t)(n / 2 + 1) * n * n * sizeof(float2)));
  w->stats = static_cast<double*>(take(PK_STATS_LEN * sizeof(double)));
  w->scatter_bytes = scatter_workspace_bytes(*g, *direct, np);
  w->scatter_ws = take(w->scatter_bytes);
  w->powspec_bytes = powspec_ws_bytes(n, nb);
  w->powspec_ws = take(w->powspec_bytes);
  w->bytes = off;
  return PK_OK;
}
}  // namespace

int pk_estimate_workspace_size(pk_plan* p, int64_t np, int stencil, const int* tile, int nb, size_t* bytes) {
  int e = check_plan(p);
  if (e) return e;
  if (!bytes) return set_error(PK_E_INVALID, "null bytes pointer");
  if (nb < 1) return set_error(PK_E_INVALID, "need at least one bin (nb=%d)", nb);
  EstimateWS w{};
  TileGeom g{};
  bool direct = false;
  e = carve_estimate(p, np, stencil, tile, nb, nullptr, &w, &g, &direct);
  if (e) return e;
  *bytes = w.bytes;
  return PK_OK;
}

int pk_estimate(pk_plan* p, const float* x, int64_t sx, const float* y, int64_t sy, const float* z, int64_t sz,
                const float* w, int64_t sw, int64_t np, float xmin, float ymin, float zmin, double box_size,
                int stencil, int wrap_mode, const int* tile, const int32_t* thr, const int32_t* jstart, int nb,
                int window_p, int flags, double shot_noise, double* sums, int64_t* counts, float* pk,
                float* nmodes, double* kmean, double* stats, void* workspace, size_t workspace_bytes,
                void* stream) {
  int e = check_plan(p);
  if (e) return e;
  e = check_bins(thr, jstart, nb, pk);
  if (e) return e;
  ScatterArgs a{};
  e = make_args(&a, x, sx, y, sy, z, sz, w, sw, np, xmin, ymin, zmin, box_size, p->n, stencil, wrap_mode);
  if (e) return e;
  EstimateWS ws{};
  TileGeom g{};
  bool direct = false;
  e = carve_estimate(p, np, stencil, tile, nb, workspace, &ws, &g, &direct);
  if (e) return e;
  if (workspace_bytes < ws.bytes) return set_error(PK_E_WORKSPACE, "estimate workspace %zu < %zu bytes", workspace_bytes, ws.bytes);
  if (window_p < 0 || window_p > 3) return set_error(PK_E_INVALID, "window exponent p=%d not in 0..3", window_p);
  hipStream_t s = as_stream(stream);
  double* st = stats ? stats : ws.stats;
  if (!ws.fused) {
    e = scatter_run(a, g, direct, nullptr, ws.mesh, st, ws.scatter_ws, ws.scatter_bytes, s);
    if (e) return e;
    return run_powspec(p, ws.mesh, st, thr, jstart, nb, window_p, flags | PK_FLAG_ZERO_DC, box_size, shot_noise,
                       ws.spec, sums, counts, pk, nmodes, kmean, ws.powspec_ws, ws.powspec_bytes, s);
  }
  // fused: sort -> deposit + z pass -> mass -> y pass -> x pass + bin -> finalize scaled by (n^3 / M)^2
  const void* sorted = nullptr;
  const int* key_start = nullptr;
  double* tile_mass = nullptr;
  const double* aux = nullptr;
  e = scatter_run(a, g, false, nullptr, nullptr, st, ws.scatter_ws, ws.scatter_bytes, s, &sorted, &key_start,
                  &tile_mass, &aux);
  if (e) return e;
  e = fused_deposit_zpass(p->n, a, g, key_start, sorted, p->tw, st, aux, ws.spec, tile_mass, s);
  if (e) return e;
  e = mass_finalize(tile_mass, g.ntiles, st, p->n, s);
  if (e) return e;
  return finish_powspec(p, ws.spec, thr, jstart, nb, window_p, flags | PK_FLAG_ZERO_DC, box_size, shot_noise,
                        nullptr, st, static_cast<double*>(ws.powspec_ws), sums, counts, pk, nmodes, kmean, s);
}

int pk_timing_enable(int on) {
  std::lock_guard<std::mutex> lk(g_tmu);
  g_timing = on != 0;
  return PK_OK;
}

int pk_timing_read(double* ms, int64_t* launches, int nstages) {
  std::vector<Rec> recs;
  {
    std::lock_guard<std::mutex> lk(g_tmu);
    recs.swap(g_recs);
  }
  for (int i = 0; i < nstages; ++i) {
    if (ms) ms[i] = 0.0;
    if (launches) launches[i] = 0;
  }
  int err = PK_OK;
  for (auto& r : recs) {
    float t = 0.0f;
    if (hipEventSynchronize(r.b) != hipSuccess || hipEventElapsedTime(&t, r.a, r.b) != hipSuccess)
      err = set_error(PK_E_LAUNCH, "event timing failed");
    if (r.stage >= 0 && r.stage < nstages) {
      if (ms) ms[r.stage] += t;
      if (launches) launches[r.stage] += 1;
    }
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  return err;
}

const char* pk_stage_name(int stage) {
  if (stage < 0 || stage >= PK_NUM_STAGES) return "";
  return g_stage_names[stage];
}

}  // extern "C"
