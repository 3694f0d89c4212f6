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
int g_timing_only = -1;   // >= 0: only this stage is timed
std::vector<Rec> g_recs;
thread_local hipEvent_t g_open[PK_NUM_STAGES] = {};
const char* g_stage_names[PK_NUM_STAGES] = {"sort_count", "scan",   "sort_place",  "deposit",
                                            "atomic",     "mass",   "zpass",       "ypass",
                                            "xpass_bin",  "finalize", "xpass_plain", "fill",
                                            "adjoint_fft", "gather",   "sort_fine"};
}  // namespace

void timing_suspend(bool on);
bool timing_suspended();

void timing_begin(int stage, hipStream_t s) {
  if (!g_timing || timing_suspended() || (g_timing_only >= 0 && stage != g_timing_only)) return;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  (void)hipEventRecord(e, s);
  g_open[stage] = e;
}

void timing_end(int stage, hipStream_t s) {
  if (!g_timing || timing_suspended() || !g_open[stage]) return;
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

// Fixed scratch of pk_powspec_mesh_noalloc: the bin partials and z-pass DC partials that
// pk_powspec_mesh keeps in its workspace.  Every device holds its own copy from module
// load until the process ends; every kernel writes the part it reads before reading it,
// so nothing needs clearing between calls.  Calls that overlap on one device share it and
// race.  Host code never passes `scratch` itself to a kernel (on the host that name is a
// stand-in, not a device address): the address comes from hipGetSymbolAddress, resolved
// once per device when a plan is created on it.
static __device__ __align__(256) unsigned char scratch[PK_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// resolves this device's scratch address (the device must be current); once per device
int resolve_noalloc_scratch(int device) {
  if (device < 0 || device >= MAX_DEVICES) return set_error(PK_E_INVALID, "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return PK_OK;
  void* ptr = nullptr;
  int e = check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "scratch symbol address");
  if (e) return e;
  if (!ptr) return set_error(PK_E_LAUNCH, "scratch symbol address is null on device %d", device);
  g_scratch_ptr[device] = ptr;
  return PK_OK;
}

void* noalloc_scratch(int device) {
  if (device < 0 || device >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

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
  a->aos = (sx == 3 && sy == 3 && sz == 3 && y == x + 1 && z == x + 2) ? 1 : 0;
  return PK_OK;
}

// workspace of pk_powspec_mesh: bin partials + z-pass DC partials (spectrum is an argument)
size_t powspec_ws_bytes(int n, int nb) {
  const int G = bin_grid(n, nb);
  return align_up((size_t)G * (size_t)(nb > 0 ? nb : 1) * BIN_PART_LEN * sizeof(double)) +
         align_up((size_t)2 * zpass_grid(n) * sizeof(double));   // z pass: DC and squares partials
}

int check_bins(const int32_t* thr, const int32_t* jstart, const int64_t* modes, const double* ksum, int nb,
               int flags, const float* pk) {
  if (nb < 1) return set_error(PK_E_INVALID, "need at least one bin (nb=%d)", nb);
  if (nb > (1 << 20)) return set_error(PK_E_INVALID, "too many bins (nb=%d)", nb);
  if (!thr || !jstart || !modes || !pk) return set_error(PK_E_INVALID, "null threshold / jstart / modes / pk pointer");
  if ((flags & PK_FLAG_KMEAN) && !ksum) return set_error(PK_E_INVALID, "PK_FLAG_KMEAN needs the ksum table");
  return PK_OK;
}

// binning fields of a job (everything the y pass, x pass + bin and finalize read)
void set_bin_fields(Job& j, const pk_plan* p, float2* spec, const int32_t* thr, const int32_t* jstart,
                    const int64_t* modes, const double* ksum, int nb, int window_p, int flags, double box,
                    double shot, double* sums, int64_t* counts, float* pk, float* nmodes, double* kmean) {
  const int n = p->n;
  j.nm = n;
  j.spec = spec;
  j.wtab = p->wtab + (size_t)window_p * n;
  j.thr = thr;
  j.jst = jstart;
  j.modes = modes;
  j.ksum = ksum;
  j.nb = nb;
  j.flags = flags;
  const double h = box / ((double)n * (double)n);
  j.norm = h * h * h;  // (box / n^2)^3, src/correlations.py:50-52
  j.shot = shot;
  j.sums = sums;
  j.counts = counts;
  j.pk = pk;
  j.nmodes = nmodes;
  j.kmean = kmean;
}

// y pass, x pass + binning and finalize on a spectrum whose z pass is done (fused: from the
// fused deposit, whose mass finalize the y pass does)
int finish_powspec(pk_plan* p, const Job& j, hipStream_t s, bool fused = false) {
  const int n = p->n;
  int e = fused ? fft_ypass_mass(n, j, p->tw, s) : fft_ypass(n, j, p->tw, s);
  if (e) return e;
  const int G = bin_grid(n, j.nb);
  e = fft_xpass_bin(n, j, p->tw, G, s);
  if (e) return e;
  return bin_finalize(j, G, s);
}

double* dc_partials(void* ws, int n, int nb) {
  return reinterpret_cast<double*>(static_cast<char*>(ws) +
                                   align_up((size_t)bin_grid(n, nb) * (size_t)nb * BIN_PART_LEN * sizeof(double)));
}

int run_powspec(pk_plan* p, const float* mesh, const double* stats, const int32_t* thr, const int32_t* jstart,
                const int64_t* modes, const double* ksum, int nb, int window_p, int flags, double box, double shot,
                float2* spec, double* sums,
                int64_t* counts, float* pk, float* nmodes, double* kmean, void* ws, size_t ws_bytes,
                hipStream_t s) {
  const int n = p->n;
  const size_t need = powspec_ws_bytes(n, nb);
  if (ws_bytes < need) return set_error(PK_E_WORKSPACE, "powspec workspace %zu < %zu bytes", ws_bytes, need);
  if (window_p < 0 || window_p > 3) return set_error(PK_E_INVALID, "window exponent p=%d not in 0..3", window_p);
  Job j{};
  j.active = 1;
  set_bin_fields(j, p, spec, thr, jstart, modes, ksum, nb, window_p, flags, box, shot, sums, counts, pk, nmodes,
                 kmean);
  j.partials = static_cast<double*>(ws);
  // with the contrast applied (ZERO_DC) the k = 0 mode is exactly 0 and no DC sum is needed
  const bool keep_dc = (flags & PK_FLAG_ZERO_DC) == 0;
  double* zp = dc_partials(ws, n, nb);   // [Gz] DC sums, then [Gz] sums of squares
  double* dc_part = keep_dc ? zp : nullptr;
  j.dc_part = dc_part;
  j.Gz = zpass_grid(n);
  j.sq_part = zp + j.Gz;
  j.sq_n = j.Gz;
  int e = fft_zpass(n, mesh, spec, p->tw, stats, dc_part, zp + j.Gz, s);
  if (e) return e;
  return finish_powspec(p, j, s);
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
  if (!e) e = resolve_noalloc_scratch(p->device);   // set-up of pk_powspec_mesh_noalloc's scratch
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
  Job j{};
  j.active = 1;
  int e = make_args(&j.a, x, sx, y, sy, z, sz, w, sw, np, xmin, ymin, zmin, box_size, n, stencil, wrap_mode);
  if (e) return e;
  if (!mesh_out) return set_error(PK_E_INVALID, "null mesh_out");
  bool direct = false;
  e = make_tile_geom(n, np, stencil, tile, &j.g, &direct);
  if (e) return e;
  j.stats = stats;
  return scatter_run(j, direct, mesh_in, mesh_out, workspace, workspace_bytes, as_stream(stream));
}

int pk_r2c(pk_plan* p, const float* mesh, const double* stats, void* spectrum, void* stream) {
  int e = check_plan(p);
  if (e) return e;
  if (!mesh || !spectrum) return set_error(PK_E_INVALID, "null mesh / spectrum");
  hipStream_t s = as_stream(stream);
  float2* spec = static_cast<float2*>(spectrum);
  e = fft_zpass(p->n, mesh, spec, p->tw, stats, nullptr, nullptr, s);
  Job j{};
  j.spec = spec;
  if (!e) e = fft_ypass(p->n, j, p->tw, s);
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
                    const int64_t* modes, const double* ksum, int nb, int window_p, int flags, double box_size,
                    double shot_noise, void* spectrum, double* sums, int64_t* counts, float* pk, float* nmodes,
                    double* kmean, void* workspace, size_t workspace_bytes, void* stream) {
  int e = check_plan(p);
  if (e) return e;
  e = check_bins(thr, jstart, modes, ksum, nb, flags, pk);
  if (e) return e;
  if (!mesh || !spectrum) return set_error(PK_E_INVALID, "null mesh / spectrum");
  if (!(box_size > 0.0)) return set_error(PK_E_INVALID, "box_size must be > 0");
  return run_powspec(p, mesh, stats, thr, jstart, modes, ksum, nb, window_p, flags, box_size, shot_noise,
                     static_cast<float2*>(spectrum), sums, counts, pk, nmodes, kmean, workspace, workspace_bytes,
                     as_stream(stream));
}

int pk_powspec_mesh_noalloc(pk_plan* p, const float* mesh, const double* stats, const int32_t* thr,
                            const int32_t* jstart, const int64_t* modes, const double* ksum, int nb, int window_p,
                            int flags, double box_size, double shot_noise, void* spectrum, double* sums,
                            int64_t* counts, float* pk, float* nmodes, double* kmean, void* stream) {
  int e = check_plan(p);
  if (e) return e;
  e = check_bins(thr, jstart, modes, ksum, nb, flags, pk);
  if (e) return e;
  if (!mesh || !spectrum) return set_error(PK_E_INVALID, "null mesh / spectrum");
  if (!(box_size > 0.0)) return set_error(PK_E_INVALID, "box_size must be > 0");
  const size_t need = powspec_ws_bytes(p->n, nb);
  if (need > PK_NOALLOC_SCRATCH_BYTES)
    return set_error(PK_E_WORKSPACE,
                     "pk_powspec_mesh_noalloc: n=%d with %d bins needs %zu bytes of scratch, above the fixed %zu "
                     "(use pk_powspec_mesh with a workspace)",
                     p->n, nb, need, (size_t)PK_NOALLOC_SCRATCH_BYTES);
  void* ws = noalloc_scratch(p->device);
  if (!ws) return set_error(PK_E_LAUNCH, "pk_powspec_mesh_noalloc: scratch address of device %d unknown", p->device);
  return run_powspec(p, mesh, stats, thr, jstart, modes, ksum, nb, window_p, flags, box_size, shot_noise,
                     static_cast<float2*>(spectrum), sums, counts, pk, nmodes, kmean, ws, need, as_stream(stream));
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
                int stencil, int wrap_mode, const int* tile, const int32_t* thr, const int32_t* jstart,
                const int64_t* modes, const double* ksum, int nb, int window_p, int flags, double shot_noise,
                double* sums, int64_t* counts, float* pk, float* nmodes, double* kmean, double* stats,
                void* workspace, size_t workspace_bytes, void* stream) {
  int e = check_plan(p);
  if (e) return e;
  e = check_bins(thr, jstart, modes, ksum, nb, flags, pk);
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
  Job j{};
  j.active = 1;
  j.a = a;
  j.g = g;
  j.stats = stats ? stats : ws.stats;
  if (!ws.fused) {
    e = scatter_run(j, direct, nullptr, ws.mesh, ws.scatter_ws, ws.scatter_bytes, s);
    if (e) return e;
    return run_powspec(p, ws.mesh, j.stats, thr, jstart, modes, ksum, nb, window_p, flags | PK_FLAG_ZERO_DC,
                       box_size, shot_noise, ws.spec, sums, counts, pk, nmodes, kmean, ws.powspec_ws,
                       ws.powspec_bytes, s);
  }
  // fused: sort -> deposit + z pass -> mass -> y pass -> x pass + bin -> finalize scaled by (n^3 / M)^2
  e = scatter_run(j, false, nullptr, nullptr, ws.scatter_ws, ws.scatter_bytes, s);
  if (e) return e;
  set_bin_fields(j, p, ws.spec, thr, jstart, modes, ksum, nb, window_p, flags | PK_FLAG_ZERO_DC, box_size,
                 shot_noise, sums, counts, pk, nmodes, kmean);
  j.partials = static_cast<double*>(ws.powspec_ws);
  j.scale_stats = j.stats;
  j.sq_part = j.aux + 2;   // the mass finalize sums the tiles' squares there
  j.sq_n = 1;
  e = fused_deposit_zpass(p->n, j, p->tw, s);
  if (e) return e;
  return finish_powspec(p, j, s, true);
}

namespace {
struct BackwardWS {
  float2* bspec;
  float* gmesh;
  float* coef;
  void* gather;
  size_t gather_bytes;
  size_t bytes;
};

BackwardWS carve_backward(int n, int64_t np, int nb, void* base) {
  BackwardWS w{};
  char* b = static_cast<char*>(base);
  size_t off = 0;
  auto take = [&](size_t bytes) -> void* {
    void* q = b ? b + off : nullptr;
    off += align_up(bytes);
    return q;
  };
  w.bspec = static_cast<float2*>(take((size_t)(n / 2 + 1) * n * n * sizeof(float2)));
  w.gmesh = static_cast<float*>(take((size_t)n * n * n * sizeof(float)));
  w.coef = static_cast<float*>(take((size_t)nb * 3 * sizeof(float)));
  w.gather_bytes = gather_workspace_bytes(np);
  w.gather = take(w.gather_bytes);
  w.bytes = off;
  return w;
}
}  // namespace

int pk_estimate_backward_workspace_size(pk_plan* p, int64_t np, int nb, size_t* bytes) {
  int e = check_plan(p);
  if (e) return e;
  if (!bytes) return set_error(PK_E_INVALID, "null bytes pointer");
  if (nb < 1) return set_error(PK_E_INVALID, "need at least one bin (nb=%d)", nb);
  if (np < 0 || np > 0x7fffffffLL) return set_error(PK_E_INVALID, "particle count out of range");
  *bytes = carve_backward(p->n, np, nb, nullptr).bytes;
  return PK_OK;
}

int pk_estimate_backward(pk_plan* p, const float* x, int64_t sx, const float* y, int64_t sy, const float* z,
                         int64_t sz, const float* w, int64_t sw, int64_t np, float xmin, float ymin, float zmin,
                         double box_size, int stencil, int wrap_mode, const int* tile, const int32_t* thr,
                         const int32_t* jstart, int nb, int window_p, const void* fwd_workspace,
                         size_t fwd_workspace_bytes, const double* stats, const double* grad_pk,
                         const int64_t* counts, float* grad_pos, float* grad_w, void* workspace,
                         size_t workspace_bytes, void* stream) {
  int e = check_plan(p);
  if (e) return e;
  if (nb < 1 || !thr || !jstart) return set_error(PK_E_INVALID, "bad bin tables");
  if (!fwd_workspace || !stats || !grad_pk || !counts || !grad_pos)
    return set_error(PK_E_INVALID, "null forward workspace / stats / grad_pk / counts / grad_pos");
  if (window_p < 0 || window_p > 3) return set_error(PK_E_INVALID, "window exponent p=%d not in 0..3", window_p);
  ScatterArgs a{};
  e = make_args(&a, x, sx, y, sy, z, sz, w, sw, np, xmin, ymin, zmin, box_size, p->n, stencil, wrap_mode);
  if (e) return e;
  EstimateWS fw{};
  TileGeom g{};
  bool direct = false;
  e = carve_estimate(p, np, stencil, tile, nb, const_cast<void*>(fwd_workspace), &fw, &g, &direct);
  if (e) return e;
  if (fwd_workspace_bytes < fw.bytes)
    return set_error(PK_E_WORKSPACE, "forward workspace %zu < %zu bytes", fwd_workspace_bytes, fw.bytes);
  BackwardWS bw = carve_backward(p->n, np, nb, workspace);
  if (workspace_bytes < bw.bytes)
    return set_error(PK_E_WORKSPACE, "backward workspace %zu < %zu bytes", workspace_bytes, bw.bytes);
  hipStream_t s = as_stream(stream);
  const int n = p->n;
  const double h = box_size / ((double)n * (double)n);
  e = adjoint_coefficients(grad_pk, counts, nb, h * h * h, bw.coef, s);
  if (e) return e;
  Job j{};
  j.active = 1;
  set_bin_fields(j, p, fw.spec, thr, jstart, counts, nullptr, nb, window_p, PK_FLAG_ZERO_DC, box_size, 0.0, nullptr,
                 nullptr, nullptr, nullptr, nullptr);
  j.scale_stats = fw.fused ? stats : nullptr;   // fused: the spectrum is of rho - c, delta_k = S spec_k
  e = fft_adjoint(n, j, p->tw, bw.coef, bw.bspec, bw.gmesh, s);
  if (e) return e;
  return gather_gradients(a, bw.gmesh, stats, grad_pos, grad_w, bw.gather, bw.gather_bytes, s);
}

// ------------------------------------------------------------------------------------
// batches
// ------------------------------------------------------------------------------------
}  // extern "C"

struct pk_batch {
  pk_plan* plan;
  int n;
  int max_jobs;
  int64_t max_np;
  int max_nb;
  int flags;
  TileGeom cap;          // capacity geometry: TSC classes, max_np particles
  size_t slot_bytes;
  char* slots;           // per-job scratch slots (caller workspace)
  Job* d_jobs;           // device job table (head of the workspace)
  Job* h_jobs[2];        // pinned host tables, used alternately
  hipEvent_t uploaded[2];
  bool pending[2];
  int cur;
  int njobs;
  BatchDims dims;
  int gmax;              // x-pass blocks of one job alone (partials rows per slot)
  // run-time tuning (pk_batch_set_tuning): counting rounds over the GPU, and the y / x
  // passes' blocks as a percentage of their default (resident-capacity) grids
  int count_rounds = 4;
  int ypass_pct = 100;
  int xpass_pct = 100;
  std::vector<Job> slot;  // scratch pointers of every slot
  struct Graph {
    BatchDims d;
    hipGraphExec_t exec;
    hipGraph_t graph;
  };
  std::vector<Graph> graphs;
  std::vector<BatchDims> seen;   // shapes already run once without capture
};

namespace pk {
namespace {
thread_local bool t_capturing = false;
}
void timing_suspend(bool on) { t_capturing = on; }
bool timing_suspended() { return t_capturing; }
}  // namespace pk

namespace {
bool same_dims(const BatchDims& a, const BatchDims& b) {
  return a.njobs == b.njobs && a.max_nblk == b.max_nblk && a.max_nkeys == b.max_nkeys && a.max_nb == b.max_nb &&
         a.lds_keys == b.lds_keys && a.n == b.n && a.max_fblk == b.max_fblk;
}

int batch_geometry(int n, int64_t max_np, const int* tile, TileGeom* cap) {
  bool direct = false;
  int e = make_tile_geom(n, max_np, PK_STENCIL_TSC, tile, cap, &direct);
  if (e) return e;
  if (direct || !fused_supported(n, *cap))
    return set_error(PK_E_UNSUPPORTED, "batches need the sorted fused path, which n=%d with %d rows per tile lacks", n,
                     cap->ty);
  return PK_OK;
}

struct SlotLayout {
  size_t spec, partials, scatter, bytes;
};

SlotLayout slot_layout(int n, int64_t max_np, int max_nb, const TileGeom& cap, int gmax) {
  SlotLayout L{};
  L.spec = 0;
  size_t off = align_up((size_t)(n / 2 + 1) * n * n * sizeof(float2));
  L.partials = off;
  off += align_up((size_t)gmax * (size_t)max_nb * BIN_PART_LEN * sizeof(double));
  L.scatter = off;
  Job cj{};
  cj.g = cap;
  cj.a.np = max_np;
  off += carve_scatter(cj, false, nullptr);
  L.bytes = align_up(off);
  return L;
}

size_t table_bytes(int max_jobs) { return align_up((size_t)max_jobs * sizeof(Job)); }

int batch_launch(pk_batch* b, hipStream_t s) {
  const int n = b->n, nj = b->njobs;
  const float2* tw = b->plan->tw;
  int e = scatter_sort_jobs(b->d_jobs, b->dims, s);
  if (e) return e;
  e = fused_deposit_zpass_jobs(n, b->cap.ty, b->d_jobs, nj, b->cap.ntiles, tw, s);
  if (e) return e;
  int gy = (int)((int64_t)ypass_jobs_grid(n, nj) * b->ypass_pct / 100);
  if (gy < 1) gy = 1;
  e = fft_ypass_jobs(n, b->d_jobs, nj, gy, tw, s);
  if (e) return e;
  int gx = (int)((int64_t)bin_jobs_grid(n, b->dims.max_nb, nj) * b->xpass_pct / 100);
  if (gx < 1) gx = 1;
  if (gx > b->gmax) gx = b->gmax;
  e = fft_xpass_bin_jobs(n, b->d_jobs, nj, b->dims.max_nb, gx, tw, s);
  if (e) return e;
  return bin_finalize_jobs(b->d_jobs, nj, b->dims.max_nb, gx, s);
}
}  // namespace

extern "C" {

int pk_batch_workspace_size(pk_plan* p, int max_jobs, int64_t max_np, int max_nb, const int* tile, size_t* bytes) {
  int e = check_plan(p);
  if (e) return e;
  if (!bytes) return set_error(PK_E_INVALID, "null bytes pointer");
  if (max_jobs < 1 || max_jobs > 65535) return set_error(PK_E_INVALID, "max_jobs=%d out of range", max_jobs);
  if (max_np < 0 || max_np > 0x7fffffffLL) return set_error(PK_E_INVALID, "max_np out of range");
  if (max_nb < 1 || max_nb > (1 << 20)) return set_error(PK_E_INVALID, "max_nb=%d out of range", max_nb);
  TileGeom cap{};
  e = batch_geometry(p->n, max_np, tile, &cap);
  if (e) return e;
  const SlotLayout L = slot_layout(p->n, max_np, max_nb, cap, bin_grid(p->n, max_nb));
  *bytes = table_bytes(max_jobs) + (size_t)max_jobs * L.bytes;
  return PK_OK;
}

int pk_batch_create(pk_plan* p, int max_jobs, int64_t max_np, int max_nb, const int* tile, int flags, void* workspace,
                    size_t workspace_bytes, pk_batch** out) {
  if (!out) return set_error(PK_E_INVALID, "null batch output pointer");
  *out = nullptr;
  size_t need = 0;
  int e = pk_batch_workspace_size(p, max_jobs, max_np, max_nb, tile, &need);
  if (e) return e;
  if (!workspace || workspace_bytes < need)
    return set_error(PK_E_WORKSPACE, "batch workspace %zu < %zu bytes", workspace_bytes, need);
  pk_batch* b = new pk_batch();
  b->plan = p;
  b->n = p->n;
  b->max_jobs = max_jobs;
  b->max_np = max_np;
  b->max_nb = max_nb;
  b->flags = flags;
  batch_geometry(p->n, max_np, tile, &b->cap);
  b->gmax = bin_grid(p->n, max_nb);
  const SlotLayout L = slot_layout(p->n, max_np, max_nb, b->cap, b->gmax);
  b->slot_bytes = L.bytes;
  b->d_jobs = static_cast<Job*>(workspace);
  b->slots = static_cast<char*>(workspace) + table_bytes(max_jobs);
  b->slot.resize(max_jobs);
  for (int i = 0; i < max_jobs; ++i) {
    char* base = b->slots + (size_t)i * L.bytes;
    Job& j = b->slot[i];
    j = Job{};
    j.g = b->cap;
    j.a.np = max_np;
    j.spec = reinterpret_cast<float2*>(base + L.spec);
    j.partials = reinterpret_cast<double*>(base + L.partials);
    carve_scatter(j, false, base + L.scatter);   // also points j.stats at the slot's own stats
  }
  for (int k = 0; k < 2; ++k) {
    b->h_jobs[k] = nullptr;
    b->uploaded[k] = nullptr;
    b->pending[k] = false;
  }
  for (int k = 0; k < 2 && !e; ++k) {
    e = check_hip(hipHostMalloc(reinterpret_cast<void**>(&b->h_jobs[k]), (size_t)max_jobs * sizeof(Job)),
                  "pinned job table");
    if (!e) e = check_hip(hipEventCreateWithFlags(&b->uploaded[k], hipEventDisableTiming), "job table event");
  }
  if (e) {
    pk_batch_destroy(b);
    return e == PK_E_LAUNCH ? PK_E_OOM : e;
  }
  b->cur = 0;
  b->njobs = 0;
  b->dims = BatchDims{0, 0, 0, 0, 0, p->n, 1};
  *out = b;
  return PK_OK;
}

int pk_batch_set_jobs(pk_batch* b, int njobs, const pk_job* jobs, void* stream) {
  if (!b) return set_error(PK_E_INVALID, "null batch");
  int e = check_plan(b->plan);
  if (e) return e;
  if (njobs < 0 || njobs > b->max_jobs) return set_error(PK_E_INVALID, "njobs=%d not in 0..%d", njobs, b->max_jobs);
  if (njobs > 0 && !jobs) return set_error(PK_E_INVALID, "null job array");
  const int k = b->cur ^ 1;   // the table not used by the latest upload
  if (b->pending[k]) {
    e = check_hip(hipEventSynchronize(b->uploaded[k]), "job table reuse");
    if (e) return e;
    b->pending[k] = false;
  }
  Job* h = b->h_jobs[k];
  BatchDims d{njobs, 1, 1, 1, 0, b->n, 1};
  // capacity grids: one LDS table size for every refill (the capacity geometry's keys if
  // they fit, else none), so captured launches repeat; jobs with more keys use HBM tables
  const int cap_lds_keys = (size_t)b->cap.nkeys <= lds_table_keys() ? b->cap.nkeys : 0;
  const int tile[3] = {1, b->cap.ty, 0};
  // counting chunk of a batch: about four rounds of blocks over the whole GPU for all its
  // particles (jobs of very different sizes balance), never more blocks per job than its slot holds
  int64_t total_np = 0;
  for (int i = 0; i < njobs; ++i) total_np += jobs[i].np > 0 ? jobs[i].np : 0;
  int dev = 0;
  (void)hipGetDevice(&dev);
  const int64_t rounds = (int64_t)b->count_rounds * num_cus(dev);
  const int64_t batch_chunk = (total_np + rounds - 1) / rounds;
  for (int i = 0; i < njobs; ++i) {
    const pk_job& q = jobs[i];
    Job j = b->slot[i];
    e = make_args(&j.a, q.x, q.sx, q.y, q.sy, q.z, q.sz, q.w, q.sw, q.np, q.xmin, q.ymin, q.zmin, q.box_size, b->n,
                  q.stencil, q.wrap_mode);
    if (!e && q.np > b->max_np) e = set_error(PK_E_INVALID, "job %d: %lld particles > batch capacity %lld", i,
                                              (long long)q.np, (long long)b->max_np);
    if (!e) e = check_bins(q.thr, q.jstart, q.modes, q.ksum, q.nb, q.flags, q.pk);
    if (!e && q.nb > b->max_nb) e = set_error(PK_E_INVALID, "job %d: %d bins > batch capacity %d", i, q.nb, b->max_nb);
    if (!e && (q.window_p < 0 || q.window_p > 3)) e = set_error(PK_E_INVALID, "job %d: window p=%d", i, q.window_p);
    bool direct = false;
    if (!e) e = make_tile_geom(b->n, q.np, q.stencil, tile, &j.g, &direct);
    if (!e && (direct || !fused_supported(b->n, j.g))) e = set_error(PK_E_UNSUPPORTED, "job %d: no fused path", i);
    if (e) return e;
    {
      int64_t c = batch_chunk;
      const int64_t floor_c = (q.np + b->cap.nblk - 1) / b->cap.nblk;   // at most cap.nblk blocks
      if (c < floor_c) c = floor_c;
      set_chunk(&j.g, q.np, c);
      if (j.g.nblk > b->cap.nblk) return set_error(PK_E_INVALID, "job %d: counting blocks exceed the slot", i);
    }
    j.active = 1;
    if (q.stats) j.stats = q.stats;
    set_bin_fields(j, b->plan, j.spec, q.thr, q.jstart, q.modes, q.ksum, q.nb, q.window_p, q.flags | PK_FLAG_ZERO_DC,
                   q.box_size, q.shot_noise, q.sums, q.counts, q.pk, q.nmodes, q.kmean);
    j.dc_part = nullptr;
    j.Gz = 0;
    j.scale_stats = j.stats;
    j.sq_part = j.aux + 2;
    j.sq_n = 1;
    h[i] = j;
    if (j.g.nblk > d.max_nblk) d.max_nblk = j.g.nblk;
    if (j.g.nkeys > d.max_nkeys) d.max_nkeys = j.g.nkeys;
    if (fine_place_blocks(q.np) > d.max_fblk) d.max_fblk = fine_place_blocks(q.np);
    if (q.nb > d.max_nb) d.max_nb = q.nb;
    if (b->flags & PK_BATCH_CAPACITY_GRIDS) {
      if (j.g.nkeys > cap_lds_keys) h[i].g.gtab = 1;
    } else if (!j.g.gtab && j.g.nkeys > d.lds_keys) {
      d.lds_keys = j.g.nkeys;
    }
  }
  if (b->flags & PK_BATCH_CAPACITY_GRIDS) {
    d.max_nblk = b->cap.nblk;
    d.max_nkeys = b->cap.nkeys;
    d.lds_keys = cap_lds_keys;
    d.max_fblk = fine_place_blocks(b->max_np);
  }
  hipStream_t s = as_stream(stream);
  if (njobs > 0) {
    e = check_hip(hipMemcpyAsync(b->d_jobs, h, (size_t)njobs * sizeof(Job), hipMemcpyHostToDevice, s),
                  "job table upload");
    if (e) return e;
    e = check_hip(hipEventRecord(b->uploaded[k], s), "job table event");
    if (e) return e;
    b->pending[k] = true;
  }
  b->cur = k;
  b->njobs = njobs;
  b->dims = d;
  return PK_OK;
}

int pk_batch_run(pk_batch* b, void* stream) {
  if (!b) return set_error(PK_E_INVALID, "null batch");
  int e = check_plan(b->plan);
  if (e) return e;
  if (b->njobs == 0) return PK_OK;
  hipStream_t s = as_stream(stream);
  bool graph = (b->flags & PK_BATCH_GRAPH) != 0 && s != nullptr;
  if (graph) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) graph = false;
  }
  if (!graph) return batch_launch(b, s);
  for (auto& g : b->graphs)
    if (same_dims(g.d, b->dims)) return check_hip(hipGraphLaunch(g.exec, s), "batch graph launch");
  bool seen = false;
  for (auto& d : b->seen) seen = seen || same_dims(d, b->dims);
  if (!seen) {   // first run of this shape: plain launches (sets kernel attributes outside capture)
    b->seen.push_back(b->dims);
    return batch_launch(b, s);
  }
  pk_batch::Graph g{b->dims, nullptr, nullptr};
  e = check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "batch capture begin");
  if (e) return e;
  timing_suspend(true);
  const int le = batch_launch(b, s);
  timing_suspend(false);
  e = check_hip(hipStreamEndCapture(s, &g.graph), "batch capture end");
  if (le) {
    if (g.graph) (void)hipGraphDestroy(g.graph);
    return le;
  }
  if (e) return e;
  e = check_hip(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0), "batch graph instantiate");
  if (e) {
    (void)hipGraphDestroy(g.graph);
    return e;
  }
  b->graphs.push_back(g);
  return check_hip(hipGraphLaunch(g.exec, s), "batch graph launch");
}

int pk_batch_set_tuning(pk_batch* b, int count_rounds, int ypass_pct, int xpass_pct) {
  if (!b) return set_error(PK_E_INVALID, "null batch");
  if (count_rounds < 1 || count_rounds > 64) return set_error(PK_E_INVALID, "count_rounds=%d not in 1..64", count_rounds);
  if (ypass_pct < 10 || ypass_pct > 400 || xpass_pct < 10 || xpass_pct > 400)
    return set_error(PK_E_INVALID, "pass grid percentages must be in 10..400");
  // captured graphs hold the old grids: drop them (the next runs capture again)
  for (auto& g : b->graphs) {
    (void)hipGraphExecDestroy(g.exec);
    (void)hipGraphDestroy(g.graph);
  }
  b->graphs.clear();
  b->seen.clear();
  b->count_rounds = count_rounds;
  b->ypass_pct = ypass_pct;
  b->xpass_pct = xpass_pct;
  return PK_OK;
}

int pk_batch_tile(pk_batch* b, int* tile) {
  if (!b || !tile) return set_error(PK_E_INVALID, "null batch / tile");
  tile[0] = 1;
  tile[1] = b->cap.ty;
  tile[2] = b->n;
  return PK_OK;
}

int pk_batch_destroy(pk_batch* b) {
  if (!b) return PK_OK;
  for (auto& g : b->graphs) {
    (void)hipGraphExecDestroy(g.exec);
    (void)hipGraphDestroy(g.graph);
  }
  for (int k = 0; k < 2; ++k) {
    if (b->pending[k] && b->uploaded[k]) (void)hipEventSynchronize(b->uploaded[k]);
    if (b->uploaded[k]) (void)hipEventDestroy(b->uploaded[k]);
    if (b->h_jobs[k]) (void)hipHostFree(b->h_jobs[k]);
  }
  delete b;
  return PK_OK;
}

int pk_timing_enable(int on) {
  std::lock_guard<std::mutex> lk(g_tmu);
  g_timing = on != 0;
  return PK_OK;
}

int pk_timing_select(int stage) {
  if (stage < -1 || stage >= PK_NUM_STAGES) return set_error(PK_E_INVALID, "stage %d not in -1..%d", stage, PK_NUM_STAGES - 1);
  std::lock_guard<std::mutex> lk(g_tmu);
  g_timing_only = stage;
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

// ------------------------------------------------------------------------------------
// seeded mock generator
// ------------------------------------------------------------------------------------
namespace {
int make_field_gen(pk_plan* p, double box, const double* kf, const double* pkf, int nk, int rayleigh, uint64_t seed,
                   void* ws, size_t ws_bytes, FieldGen* g, hipStream_t s) {
  int e = check_plan(p);
  if (e) return e;
  if (!(box > 0.0)) return set_error(PK_E_INVALID, "box_size must be > 0");
  if (!kf || !pkf) return set_error(PK_E_INVALID, "null kf / Pkf table");
  if (nk < 2) return set_error(PK_E_INVALID, "the P(k) table needs at least 2 entries (nk=%d)", nk);
  if (rayleigh != 0 && rayleigh != 1) return set_error(PK_E_INVALID, "Rayleigh_sampling must be 0 or 1");
  const size_t need = align_up(gauss_table_len(p->n) * sizeof(float));
  if (!ws || ws_bytes < need) return set_error(PK_E_WORKSPACE, "Gaussian-field workspace %zu < %zu bytes", ws_bytes, need);
  float* amp = static_cast<float*>(ws);
  e = gauss_table(p->n, box, kf, pkf, nk, amp, s);
  if (e) return e;
  g->amp = amp;
  g->k0 = (uint32_t)seed;
  g->k1 = (uint32_t)(seed >> 32);
  g->n = p->n;
  g->rayleigh = rayleigh;
  g->l2pi = (float)(box / (2.0 * M_PI));
  return PK_OK;
}

int check_pop(const float* field, int n, double box) {
  if (!field) return set_error(PK_E_INVALID, "null field");
  if (n < 1 || n > 1024) return set_error(PK_E_INVALID, "mesh side n=%d out of range (1..1024)", n);
  if (!(box > 0.0)) return set_error(PK_E_INVALID, "box_size must be > 0");
  return PK_OK;
}
}  // namespace

extern "C" {

int pk_gaussian_workspace_size(int n, size_t* bytes) {
  if (!bytes) return set_error(PK_E_INVALID, "null bytes pointer");
  int e = check_n(n);
  if (e) return e;
  *bytes = align_up(gauss_table_len(n) * sizeof(float));
  return PK_OK;
}

int pk_gaussian_field(pk_plan* p, double box_size, const double* kf, const double* pkf, int nk, int rayleigh,
                      uint64_t seed, void* spectrum, void* workspace, size_t workspace_bytes, void* stream) {
  hipStream_t s = as_stream(stream);
  FieldGen g{};
  int e = make_field_gen(p, box_size, kf, pkf, nk, rayleigh, seed, workspace, workspace_bytes, &g, s);
  if (e) return e;
  if (!spectrum) return set_error(PK_E_INVALID, "null spectrum");
  return gauss_spectrum(g, static_cast<float2*>(spectrum), s);
}

int pk_c2r(pk_plan* p, const void* spectrum, void* scratch, float* mesh, void* stream) {
  int e = check_plan(p);
  if (e) return e;
  if (!spectrum || !scratch || !mesh) return set_error(PK_E_INVALID, "null spectrum / scratch / mesh");
  FieldGen g{};
  g.n = p->n;
  return fft_c2r(p->n, C2R_SPECTRUM, static_cast<const float2*>(spectrum), static_cast<float2*>(scratch), g, p->tw,
                 mesh, as_stream(stream));
}

int pk_gaussian_mesh(pk_plan* p, double box_size, const double* kf, const double* pkf, int nk, int rayleigh,
                     uint64_t seed, int component, void* scratch, float* mesh, void* workspace, size_t workspace_bytes,
                     void* stream) {
  if (component != PK_FIELD_DELTA && component != PK_FIELD_PSI_Z)
    return set_error(PK_E_INVALID, "unknown field component %d", component);
  if (!scratch || !mesh) return set_error(PK_E_INVALID, "null scratch / mesh");
  hipStream_t s = as_stream(stream);
  FieldGen g{};
  int e = make_field_gen(p, box_size, kf, pkf, nk, rayleigh, seed, workspace, workspace_bytes, &g, s);
  if (e) return e;
  return fft_c2r(p->n, component == PK_FIELD_DELTA ? C2R_DELTA : C2R_PSI_Z, nullptr, static_cast<float2*>(scratch), g,
                 p->tw, mesh, s);
}

int pk_populate_workspace_size(int n, size_t* bytes) {
  if (!bytes) return set_error(PK_E_INVALID, "null bytes pointer");
  if (n < 1 || n > 1024) return set_error(PK_E_INVALID, "mesh side n=%d out of range (1..1024)", n);
  *bytes = populate_ws_bytes(n);
  return PK_OK;
}

int pk_populate_count(const float* field, int n, double box_size, double density, double bias, uint64_t seed,
                      void* workspace, size_t workspace_bytes, int64_t* total, void* stream) {
  int e = check_pop(field, n, box_size);
  if (e) return e;
  if (!(density >= 0.0)) return set_error(PK_E_INVALID, "density must be >= 0");
  const size_t need = populate_ws_bytes(n);
  if (!workspace || workspace_bytes < need)
    return set_error(PK_E_WORKSPACE, "population workspace %zu < %zu bytes", workspace_bytes, need);
  return populate_count(field, n, box_size, density, bias, seed, workspace, total, as_stream(stream));
}

int pk_populate_place(const float* field, const float* psi_z, double growth_rate, int n, double box_size, double bias,
                      uint64_t seed, void* workspace, size_t workspace_bytes, float* pos, int64_t capacity,
                      void* stream) {
  int e = check_pop(field, n, box_size);
  if (e) return e;
  if (capacity < 0) return set_error(PK_E_INVALID, "negative capacity");
  if (capacity > 0 && !pos) return set_error(PK_E_INVALID, "null positions");
  const size_t need = populate_ws_bytes(n);
  if (!workspace || workspace_bytes < need)
    return set_error(PK_E_WORKSPACE, "population workspace %zu < %zu bytes", workspace_bytes, need);
  return populate_place(field, psi_z, growth_rate, n, box_size, bias, seed, workspace, pos, capacity,
                        as_stream(stream));
}

}  // extern "C"

// ------------------------------------------------------------------------------------
// configuration-space multipoles
// ------------------------------------------------------------------------------------
namespace {
size_t xi_ws_bytes(int n, int nb) {
  const size_t half = (size_t)(n / 2 + 1) * n * n * sizeof(float2);
  return align_up(half) + align_up((size_t)xi_grid_x(n) * sizeof(double)) + align_up(2 * sizeof(double)) +
         align_up((size_t)nb * 3 * sizeof(unsigned long long)) + align_up((size_t)nb * sizeof(int));
}
}  // namespace

extern "C" {

int pk_xi_workspace_size(pk_plan* p, int nb, size_t* bytes) {
  int e = check_plan(p);
  if (e) return e;
  if (!bytes) return set_error(PK_E_INVALID, "null bytes pointer");
  if (nb < 1 || nb > PK_XI_MAX_BINS) return set_error(PK_E_INVALID, "xi needs 1..%d bins (nb=%d)", PK_XI_MAX_BINS, nb);
  *bytes = xi_ws_bytes(p->n, nb);
  return PK_OK;
}

int pk_xi_mesh(pk_plan* p, const float* mesh, const int32_t* thr, const int32_t* jstart, const int64_t* modes, int nb,
               int window_p, int flags, double* sums, float* xi, float* nmodes, void* workspace,
               size_t workspace_bytes, void* stream) {
  int e = check_plan(p);
  if (e) return e;
  const int n = p->n;
  if (!mesh || !thr || !jstart || !modes || !sums || !xi || !nmodes)
    return set_error(PK_E_INVALID, "null mesh / table / output pointer");
  if (nb < 1 || nb > PK_XI_MAX_BINS) return set_error(PK_E_INVALID, "xi needs 1..%d bins (nb=%d)", PK_XI_MAX_BINS, nb);
  if (window_p < 0 || window_p > 3) return set_error(PK_E_INVALID, "window exponent p=%d not in 0..3", window_p);
  const size_t need = xi_ws_bytes(n, nb);
  if (!workspace || workspace_bytes < need)
    return set_error(PK_E_WORKSPACE, "xi workspace %zu < %zu bytes", workspace_bytes, need);
  char* w = static_cast<char*>(workspace);
  float2* spec = reinterpret_cast<float2*>(w);
  w += align_up((size_t)(n / 2 + 1) * n * n * sizeof(float2));
  XiArgs a{};
  a.xi0_part = reinterpret_cast<double*>(w);
  a.gx = xi_grid_x(n);
  w += align_up((size_t)a.gx * sizeof(double));
  a.scale = reinterpret_cast<double*>(w);
  w += align_up(2 * sizeof(double));
  a.fixed = reinterpret_cast<unsigned long long*>(w);
  w += align_up((size_t)nb * 3 * sizeof(unsigned long long));
  a.nanflag = reinterpret_cast<int*>(w);
  a.thr = thr;
  a.jst = jstart;
  a.modes = modes;
  a.nb = nb;
  a.jlen = isqrt_c(3 * (n / 2) * (n / 2)) + 2;
  a.empty_nan = (flags & PK_XI_EMPTY_NAN) ? 1 : 0;
  a.sums = sums;
  a.xi = xi;
  a.nmodes = nmodes;
  return fft_xi(n, mesh, spec, p->tw, p->wtab + (size_t)window_p * n, a, as_stream(stream));
}

}  // extern "C"

// ------------------------------------------------------------------------------------
// bispectrum
// ------------------------------------------------------------------------------------
namespace {
size_t bispec_ws_bytes(int n, int nshell) {
  const size_t half = align_up((size_t)(n / 2 + 1) * n * n * sizeof(float2));
  const size_t mesh = align_up((size_t)n * n * n * sizeof(float));
  return 3 * half + 4 * mesh + align_up((size_t)nshell * BK_SLOTS * BK_PART_STRIDE * sizeof(double));
}
}  // namespace

extern "C" {

int pk_bispec_workspace_size(pk_plan* p, int nshell, size_t* bytes) {
  int e = check_plan(p);
  if (e) return e;
  if (!bytes) return set_error(PK_E_INVALID, "null bytes pointer");
  if (nshell < 2 || nshell > PK_BISPEC_MAX_SHELLS)
    return set_error(PK_E_INVALID, "bispectrum needs 2..%d shells (got %d)", PK_BISPEC_MAX_SHELLS, nshell);
  *bytes = bispec_ws_bytes(p->n, nshell);
  return PK_OK;
}

int pk_bispec_mesh(pk_plan* p, const float* mesh, const int32_t* shell_lo, const int32_t* shell_hi, int nshell,
                   int window_p, double box_size, const double* i_sums, double* sums, float* pk, float* bk, float* qk,
                   void* workspace, size_t workspace_bytes, void* stream) {
  int e = check_plan(p);
  if (e) return e;
  const int n = p->n;
  if (!mesh || !shell_lo || !shell_hi || !sums || !pk) return set_error(PK_E_INVALID, "null mesh / shell / output");
  if (nshell < 2 || nshell > PK_BISPEC_MAX_SHELLS)
    return set_error(PK_E_INVALID, "bispectrum needs 2..%d shells (got %d)", PK_BISPEC_MAX_SHELLS, nshell);
  if (nshell > 2 && (!bk || !qk)) return set_error(PK_E_INVALID, "null B / Q output");
  if (window_p < 0 || window_p > 3) return set_error(PK_E_INVALID, "window exponent p=%d not in 0..3", window_p);
  if (!(box_size > 0.0)) return set_error(PK_E_INVALID, "box_size must be > 0");
  const size_t need = bispec_ws_bytes(n, nshell);
  if (!workspace || workspace_bytes < need)
    return set_error(PK_E_WORKSPACE, "bispectrum workspace %zu < %zu bytes", workspace_bytes, need);
  std::vector<int> lo(shell_lo, shell_lo + nshell), hi(shell_hi, shell_hi + nshell);
  char* w = static_cast<char*>(workspace);
  const size_t half = align_up((size_t)(n / 2 + 1) * n * n * sizeof(float2));
  const size_t msz = align_up((size_t)n * n * n * sizeof(float));
  float2* spec = reinterpret_cast<float2*>(w);
  float2* scratch = reinterpret_cast<float2*>(w + half);
  float2* qspec = reinterpret_cast<float2*>(w + 2 * half);
  float* meshes[4];
  for (int i = 0; i < 4; ++i) meshes[i] = reinterpret_cast<float*>(w + 3 * half + i * msz);
  BispecArgs a{};
  a.nshell = nshell;
  a.box = box_size;
  a.part = reinterpret_cast<double*>(w + 3 * half + 4 * msz);
  a.i_sums = i_sums;
  a.sums = sums;
  a.pk = pk;
  a.bk = bk;
  a.qk = qk;
  return fft_bispec(n, mesh, spec, scratch, qspec, meshes, p->tw, p->wtab + (size_t)window_p * n, lo.data(), hi.data(),
                    a, as_stream(stream));
}

}  // extern "C"
