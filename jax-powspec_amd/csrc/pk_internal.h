// Internal (non-exported) interfaces shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace pk {

int num_cus(int device);

// ---- FFT passes (pk_fft.hip) --------------------------------------------------------
bool fft_supported(int n);
int zpass_grid(int n);
int fft_zpass(int n, const float* mesh, float2* spec, const float2* tw, const double* stats, double* dc_part,
              hipStream_t s);
int fft_ypass(int n, float2* spec, const float2* tw, hipStream_t s);
int fft_xpass(int n, float2* spec, const float2* tw, hipStream_t s);
int bin_grid(int n, int nb);
int fft_xpass_bin(int n, const float2* spec, const float2* tw, const float* wtab, const int* thr, const int* jst,
                  int nb, int jmax, int flags, double* partials, int grid, hipStream_t s);
int bin_finalize(const double* partials, int G, int nb, const int* thr, const double* dc_part, int Gz, double norm,
                 double shot, double* sums, int64_t* counts, float* pk, float* nmodes, double* kmean, hipStream_t s);

// ---- mass assignment (pk_scatter.hip) -----------------------------------------------
struct ScatterArgs {
  const float* x;
  const float* y;
  const float* z;
  const float* w;
  int64_t sx, sy, sz, sw;
  int64_t np;
  float xmin, ymin, zmin, inv;  // inv = 1 / ((float)box / (float)n), float32 as the reference
  int n;
  int stencil;
  int wrap_mode;
};

struct TileGeom {
  int td[3];    // tile dims (cells)
  int nt[3];    // tiles per axis
  int ncls;     // classes per tile: 8 (CIC variants) or 27 (TSC)
  int ntiles;
  int nkeys;    // ntiles * ncls + 1 (last key = irregular particles)
};

// validates/derives the tile geometry; tile == nullptr -> default; returns PK_OK or error
int make_tile_geom(int n, int stencil, const int* tile, TileGeom* g, bool* direct);
size_t scatter_workspace_bytes(const TileGeom& g, bool direct, int64_t np);
int scatter_run(const ScatterArgs& a, const TileGeom& g, bool direct, const float* mesh_in, float* mesh_out,
                double* stats, void* ws, size_t ws_bytes, hipStream_t s);

inline size_t align_up(size_t v, size_t a = 256) { return (v + a - 1) / a * a; }

}  // namespace pk
