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
int bin_finalize(const double* partials, int G, int nb, const int* thr, const double* dc_part, int Gz,
                 const double* scale_stats, double norm, double shot, double* sums, int64_t* counts, float* pk,
                 float* nmodes, double* kmean, hipStream_t s);

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

// Row-bundle tiles of the sorted deposit: tile (xt, yt) = cells x = xt,
// y in [yt*ty, (yt+1)*ty), all z (full rows, so no tile boundary along z).
struct TileGeom {
  int ty;       // rows per tile (divides n, >= 2 unless n < 2)
  int nyt;      // n / ty
  int ntiles;   // n * nyt
  int ncls;     // y classes: 2 (CIC variants) or 3 (TSC)
  int nkeys;    // ntiles * ncls + 1 (last key = irregular particles)
  int chunk;    // particles per counting block (<= 65535: 16-bit LDS counters)
  int nblk;     // counting blocks
  int fixed;    // 1: deposit into 64-bit fixed-point LDS cells (ty * n * 8 bytes fit), 0: float
};

// LDS bytes up to which the deposit accumulates in 64-bit fixed point: gfx950 LDS
// integer atomics run ~13x faster than float ones, and integer sums are order-independent
constexpr size_t FIXED_LDS_MAX = 131072;
// fixed-point scale (aux[0] = S, aux[1] = 1 / S) lives in the scatter workspace; deposits
// are rint(weight * S) with S = 2^32 / 2^ceil(log2 max|w|), so |sum| < 2^63 for < 2^31 particles
constexpr int AUX_LEN = 4;

// validates/derives the tile geometry; tile == nullptr -> default rows per tile;
// tile[0] == 0 -> direct global atomics; otherwise tile[1] = rows per tile
int make_tile_geom(int n, int64_t np, int stencil, const int* tile, TileGeom* g, bool* direct);
// fused assignment + z pass (pk_fused.hip): deposit into LDS rows, subtract stats[3],
// R2C along z, write the spectrum; per-tile FP64 masses into tile_mass
int fused_deposit_zpass(int n, const ScatterArgs& a, const TileGeom& g, const int* key_start,
                        const void* sorted, const float2* tw, const double* stats, const double* aux,
                        float2* spec, double* tile_mass, hipStream_t s);
bool fused_supported(int n, const TileGeom& g);
size_t scatter_workspace_bytes(const TileGeom& g, bool direct, int64_t np);
// mesh_out == nullptr (tiled mode only): stop after sorting; the caller runs the fused
// deposit + z pass on the sorted records (sorted_out, key_start_out, tile_mass_out)
int scatter_run(const ScatterArgs& a, const TileGeom& g, bool direct, const float* mesh_in, float* mesh_out,
                double* stats, void* ws, size_t ws_bytes, hipStream_t s, const void** sorted_out = nullptr,
                const int** key_start_out = nullptr, double** tile_mass_out = nullptr,
                const double** aux_out = nullptr);
int mass_finalize(const double* tile_mass, int ntiles, double* stats, int n, hipStream_t s);

inline size_t align_up(size_t v, size_t a = 256) { return (v + a - 1) / a * a; }

}  // namespace pk
