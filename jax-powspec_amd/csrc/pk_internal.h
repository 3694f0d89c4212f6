// Internal (non-exported) interfaces shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace pk {

int num_cus(int device);

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
  int aos;      // x, y, z interleaved (y = x + 1, z = x + 2, strides 3): one 12-byte load
};

// Row-bundle tiles of the sorted deposit: tile (xt, yt) = cells x = xt,
// y in [yt*ty, (yt+1)*ty), all z (full rows, so no tile boundary along z).
struct TileGeom {
  int ty;       // rows per tile (divides n, >= 2 unless n < 2)
  int nyt;      // n / ty
  int ntiles;   // n * nyt
  int ncls;     // y classes: 2 (CIC variants) or 3 (TSC)
  int nkeys;    // ntiles * ncls + 1 (last key = irregular particles)
  int chunk;    // particles per counting block (<= 65535: packed 16-bit LDS counters)
  int nblk;     // counting blocks
  int fixed;    // 1: deposit into 64-bit fixed-point LDS cells (ty * n * 8 bytes fit), 0: float
  unsigned ty_mul;   // ceil(2^32 / ty): y / ty = umulhi(y, ty_mul) for 0 <= y < n
  int gtab;     // 1: the count and place kernels keep their per-key tables in HBM (the
                // block's row of the count matrix, global atomics): nkeys too large for LDS
};

// LDS bytes up to which the deposit accumulates in 64-bit fixed point: gfx950 LDS
// integer atomics run ~13x faster than float ones, and integer sums are order-independent
constexpr size_t FIXED_LDS_MAX = 131072;
// fixed-point scale (aux[0] = S, aux[1] = 1 / S) lives in the scatter workspace; deposits
// are rint(weight * S) with S = 2^30 / 2^e, max|w| < 2^e, so |sum| < 2^61 for < 2^31 particles
constexpr int AUX_LEN = 4;

// One estimate's arguments and scratch as the shared kernels see them.  The plain entry
// points pass a Job by value (kernel argument); batched launches read jobs[blockIdx.y]
// from a device table, so one launch serves many catalogues of the same mesh side.
struct Job {
  ScatterArgs a;
  TileGeom g;
  int active;            // 0: every kernel skips this job (unused batch slot)
  // sorted-deposit scratch
  unsigned* mat;         // [nblk][nkeys] counts -> offsets inside each key
  int* totals;           // [nkeys]
  int* key_start;        // [nkeys + 1]
  int* bsum;             // scan block sums
  double* wsum;          // [nblk] weight sums
  double* wmax;          // [nblk] max |w|
  double* aux;           // [AUX_LEN] fixed-point scale
  void* sorted;          // records
  void* staged;          // records grouped by x plane only (the first placing pass)
  unsigned* koff;        // [nkeys] records placed so far in each key (second pass)
  double* tile_mass;     // [ntiles]
  double* tile_sq;       // [ntiles] fused tiles' sums of (rho - c)^2 (nullptr: not reduced)
  double* stats;         // [PK_STATS_LEN]
  // spectrum and binning
  float2* spec;                // [n/2+1][n][n]
  const float* wtab;           // window table of this job's exponent
  const int* thr;              // [nb + 1]
  const int* jst;              // isqrt start table
  const int64_t* modes;        // [nb] exact mode counts
  const double* ksum;          // [nb] sums of sqrt(m) (nullable without PK_FLAG_KMEAN)
  int nb;
  int flags;
  double* partials;            // [G][nb][BIN_PART_LEN] x-pass block partials
  const double* dc_part;       // FP64 k = 0 partials of the z pass (nullable)
  int Gz;                      // number of dc_part entries
  // FP64 partial sums of the squared real field behind the spectrum, summed in a fixed
  // order: n^3 times their total bounds the binned power (Parseval), which fixes the
  // exponent of the bins' 64-bit fixed-point sums
  const double* sq_part;
  int sq_n;
  int nm;                      // mesh side of the spectrum
  const double* scale_stats;   // sums scaled by scale_stats[1]^2 (fused contrast), nullable
  double norm;                 // (box / n^2)^3
  double shot;
  double* sums;
  int64_t* counts;
  float* pk;
  float* nmodes;
  double* kmean;
};

// ---- FFT passes and binning (pk_fft.hip) -------------------------------------------------
// per-block partials of the binning x pass: [G][nb][BIN_PART_LEN] doubles holding S_0, S_2,
// S_4 and the number of stored modes the block binned (an exact integer)
constexpr int BIN_PART_LEN = 4;
bool fft_supported(int n);
int zpass_grid(int n);
int fft_zpass(int n, const float* mesh, float2* spec, const float2* tw, const double* stats, double* dc_part,
              double* sq_part,
              hipStream_t s);
int fft_xpass(int n, float2* spec, const float2* tw, hipStream_t s);
int bin_grid(int n, int nb);   // x-pass blocks of one estimate
// blocks per job of a batch (y pass; x pass + binning, at most bin_grid(n, nb))
int ypass_jobs_grid(int n, int njobs);
int bin_jobs_grid(int n, int nb, int njobs);
// single job (by value) or a device table of njobs (then gy / gx = blocks per job)
int fft_ypass(int n, const Job& j, const float2* tw, hipStream_t s);
// y pass of a fused estimate, whose last block also does the mass finalize (stats[0],
// stats[1], aux[2] from the tiles' masses and squares); the batched y pass does the same
// for every job
int fft_ypass_mass(int n, const Job& j, const float2* tw, hipStream_t s);
int fft_xpass_bin(int n, const Job& j, const float2* tw, int grid, hipStream_t s);
int bin_finalize(const Job& j, int G, hipStream_t s);
// backward of the binning: per-bin coefficients coef[nb][3] (float) and j's forward
// spectrum (after the y pass) -> dLoss/d delta on the mesh gmesh [n][n][n]; bspec is a
// half-spectrum scratch buffer
int fft_adjoint(int n, const Job& j, const float2* tw, const float* coef, float2* bspec, float* gmesh, hipStream_t s);
int ypass_grid(int n);
int fft_ypass_jobs(int n, const Job* jobs, int njobs, int gy, const float2* tw, hipStream_t s);
int fft_xpass_bin_jobs(int n, const Job* jobs, int njobs, int max_nb, int gx, const float2* tw, hipStream_t s);
int bin_finalize_jobs(const Job* jobs, int njobs, int max_nb, int G, hipStream_t s);

// ---- mass assignment (pk_scatter.hip) ------------------------------------------------------
// validates/derives the tile geometry; tile == nullptr -> default rows per tile;
// tile[0] == 0 -> direct global atomics; otherwise tile[1] = rows per tile
int make_tile_geom(int n, int64_t np, int stencil, const int* tile, TileGeom* g, bool* direct);
// most keys whose count tables fit LDS (more: HBM tables)
size_t lds_table_keys();
// blocks of the second placing pass for np particles
int fine_place_blocks(int64_t np);
// particles per counting block (clamped to the supported range) and the block count
void set_chunk(TileGeom* g, int64_t np, int64_t chunk);
size_t scatter_workspace_bytes(const TileGeom& g, bool direct, int64_t np);
// points the sorted-deposit scratch of j into base (nullptr: sizes only); returns bytes
size_t carve_scatter(Job& j, bool direct, void* base);
// mesh_out == nullptr (tiled mode only): stop after sorting; the caller runs the fused
// deposit + z pass on j's sorted records
int scatter_run(Job& j, bool direct, const float* mesh_in, float* mesh_out, void* ws, size_t ws_bytes,
                hipStream_t s);
int mass_finalize(const Job& j, hipStream_t s);
// batched sort (count, scan, place) and mass finalize over a device job table; the
// bounds are maxima over the jobs (blocks beyond a job's own extent return at once)
struct BatchDims {
  int njobs;
  int max_nblk;
  int max_nkeys;
  int max_nb;
  int lds_keys;   // per-key LDS tables of count sized for this many keys (jobs with more
                  // keys use HBM tables, gtab)
  int n;          // mesh side of every job
  int max_fblk;   // blocks of the second placing pass of the largest job
};
int scatter_sort_jobs(const Job* jobs, const BatchDims& d, hipStream_t s);

// ---- backward of the assignment (pk_adjoint.hip) ------------------------------------------
// coef[b][l] = grad_pk[b][l] norm (2l+1) / counts[b] (0 for empty bins)
int adjoint_coefficients(const double* grad_pk, const int64_t* counts, int nb, double norm, float* coef,
                         hipStream_t s);
size_t gather_workspace_bytes(int64_t np);
// dLoss/d position (grad_pos [np][3], physical units) and dLoss/d weight (nullable) from
// g = dLoss/d delta on the mesh, through delta = rho n^3 / M - 1 (stats: forward M, n^3 / M)
int gather_gradients(const ScatterArgs& a, const float* gmesh, const double* stats, float* grad_pos, float* grad_w,
                     void* ws, size_t ws_bytes, hipStream_t s);

// ---- fused assignment + z pass (pk_fused.hip) ----------------------------------------------
// deposit into LDS rows, subtract stats[3], R2C along z, write the spectrum; per-tile
// FP64 masses into tile_mass.  Row-bundle heights with a compiled kernel: fused_ty_ok.
bool fused_supported(int n, const TileGeom& g);
int fused_deposit_zpass(int n, const Job& j, const float2* tw, hipStream_t s);
int fused_deposit_zpass_jobs(int n, int ty, const Job* jobs, int njobs, int ntiles, const float2* tw, hipStream_t s);

// ---- seeded mock generator (pk_mock.hip; the inverse x pass in pk_fft.hip) ---------------
// Gaussian half-spectrum modes: amp[m] = sqrt(P (n^2/L)^3) at |k| = k_F sqrt(m),
// m = 0 .. 3 (n/2)^2; Philox key (k0, k1) from the seed
struct FieldGen {
  const float* amp;
  uint32_t k0, k1;
  int n;
  int rayleigh;
  float l2pi;   // L / (2 pi): physical units of the displacement component
};
// sources of the inverse transform: a stored half spectrum, the generated Gaussian field
// or its z displacement, a k-shell of a stored spectrum times the window, the shell itself
enum { C2R_SPECTRUM = 0, C2R_DELTA = 1, C2R_PSI_Z = 2, C2R_SHELL = 3, C2R_SHELL_ONE = 4 };
// k-shell lo <= m < hi (m = kx^2 + ky^2 + kz^2); wt: per-axis window row of the plan
struct ShellMask {
  int lo, hi;
  const float* wt;
};
// numpy.fft.irfftn of a half spectrum ([kz][kx][ky]; src == C2R_SPECTRUM, read from
// spec_in, which may equal scratch) or of the generated delta / psi_z modes; scratch is a
// half-spectrum buffer, mesh [n][n][n] float32
int fft_c2r(int n, int src, const float2* spec_in, float2* scratch, const FieldGen& g, const float2* tw, float* mesh,
            hipStream_t s);

// amplitude table length 3 (n/2)^2 + 1 and its fill from (kf, pkf) (FP64 device arrays)
size_t gauss_table_len(int n);
int gauss_table(int n, double box, const double* kf, const double* pkf, int nk, float* amp, hipStream_t s);
// the half spectrum [kz][kx][ky] of the generated field
int gauss_spectrum(const FieldGen& g, float2* spec, hipStream_t s);
// Poisson population of a field (rho = exp(bias * field), or the field when bias = 0):
// count writes per-block offsets and the rate scale into ws and the total into *total;
// place (after count, same ws and arguments) writes particles [total][3]
size_t populate_ws_bytes(int n);
int populate_count(const float* field, int n, double box, double density, double bias, uint64_t seed, void* ws,
                   int64_t* total, hipStream_t s);
int populate_place(const float* field, const float* psi, double growth, int n, double box, double bias, uint64_t seed,
                   void* ws, float* pos, int64_t capacity, hipStream_t s);

// ---- bispectrum (pk_fft.hip) -----------------------------------------------------------------
// reduction slots per shell: sum delta_i^2, sum I_i^2, sum delta1 delta2 delta_i, sum I1 I2 I_i
constexpr int BK_SLOTS = 4;
constexpr int BK_PART_STRIDE = 4096;   // per-slot partial capacity (>= any reduction grid)
struct BispecArgs {
  int nshell;
  double box;
  double* part;          // [nshell][BK_SLOTS][BK_PART_STRIDE]
  const double* i_sums;  // nullable [nshell][2]: known sum I_i^2, sum I1 I2 I_i (no I transforms)
  double* sums;          // [nshell][BK_SLOTS] (output)
  float* pk;             // [nshell]
  float* bk;             // [nshell - 2]
  float* qk;             // [nshell - 2]
};
// thresholds of up to 8 shells, passed by value to the Fourier-space shell sums
struct ShellChunk {
  int lo[8];
  int hi[8];
};
// mesh -> delta_k (r2c into spec) -> the first two shells' fields (masked inverses through
// scratch) and their products -> spectra of the products (scratch, qspec) -> later shells'
// sums in Fourier space -> finalize; meshes m[4] = delta1, delta2, I1, I2 (n^3 float32 each);
// lo / hi: host thresholds
int fft_bispec(int n, const float* mesh, float2* spec, float2* scratch, float2* qspec, float* const* meshes,
               const float2* tw, const float* wt, const int* lo, const int* hi, const BispecArgs& a, hipStream_t s);

// ---- configuration-space multipoles (pk_fft.hip) -------------------------------------------
// largest bin count of the separation binning (its 64-bit accumulators live in LDS)
constexpr int XI_MAX_BINS = 1024;   // = PK_XI_MAX_BINS
struct XiArgs {
  const int* thr;                 // [nb + 1] integer thresholds on m = sx^2 + sy^2 + sz^2
  const int* jst;                 // isqrt start table
  const int64_t* modes;           // [nb] cells of the full grid per bin
  int nb;
  int jlen;                       // entries of jst
  int empty_nan;                  // empty bins: NaN and count 0 (else 0 and count inf)
  double* xi0_part;               // [Gx] FP64 partials of xi(0) (x pass)
  int gx;
  double* scale;                  // [2] fixed-point scale S = 2^e and 1 / S
  unsigned long long* fixed;      // [nb][3] fixed-point bin sums (two's complement)
  int* nanflag;                   // [nb] bin holds s = 0 (mu undefined: l = 2, 4 NaN)
  double* sums;                   // [nb][3] FP64 bin sums (output)
  float* xi;                      // [nb][3] (output)
  float* nmodes;                  // [nb] (output)
};
int xi_grid_x(int n);
// mesh -> forward z, y passes -> x pass |delta_k|^2 W + inverse -> inverse y -> C2R rows
// binned by separation -> finalize; spec is a half-spectrum buffer, wt the window row of
// the plan (exponent p)
int fft_xi(int n, const float* mesh, float2* spec, const float2* tw, const float* wt, const XiArgs& a,
           hipStream_t s);

inline size_t align_up(size_t v, size_t a = 256) { return (v + a - 1) / a * a; }

}  // namespace pk
