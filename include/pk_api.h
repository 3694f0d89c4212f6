/*
 * pk_api.h -- C ABI of the MI355X (gfx950) power-spectrum hot path.
 *
 * Path: particles -> CIC/TSC mass assignment -> density contrast -> 3D R2C FFT ->
 *       window deconvolution -> |delta_k|^2 k-shell binning with Legendre l = 0, 2, 4
 *       weights -> P_l(k) (+ shot-noise subtraction).
 * On the same plans and transforms: configuration-space multipoles xi_l(s), the FFT
 * bispectrum, and a seeded lognormal mock generator (Gaussian field, inverse FFT,
 * Poisson population).
 *
 * The reference (dforero0896/jax-powspec) is pure Python/JAX and has no FFI; each entry
 * point below replaces the reference function named in its comment.  The Python drop-in
 * (jax_powspec_amd.mas / jax_powspec_amd.correlations) binds these symbols with ctypes.
 *
 * Conventions
 *   - every function returns int status: PK_OK (0) or a negative PK_E_* code; the message
 *     of the last failure on the calling thread is returned by pk_last_error().
 *   - all array pointers are DEVICE pointers (HBM) unless the comment says "host";
 *     the library never frees caller memory.
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream); every launch
 *     is stream-ordered and asynchronous; nothing synchronises the host except
 *     pk_timing_read().
 *   - a pk_plan owns only small device tables (twiddles, window tables) for one mesh side n
 *     on the device that was current at creation; scratch memory is caller-provided
 *     (query the size with the *_workspace_size functions).
 *   - a plan may be used from one stream at a time; use one plan per concurrent stream.
 *
 * Layouts
 *   - mesh: float32 [n][n][n], C order (x slowest, z fastest) -- the reference's delta.
 *   - spectrum: complex64 (float2) [n/2+1][n][n] indexed [kz][kx][ky] (ky fastest).
 *     Entry (kz, i, j) equals numpy.fft.rfftn(mesh)[i, j, kz].
 *   - bin thresholds: int32 thr[nb+1]; bin b holds modes with thr[b] <= m < thr[b+1] where
 *     m = kx^2 + ky^2 + kz^2 (integer wavenumbers), except that thr[nb] is the first m above
 *     the closed last edge.  jstart: int32[jmax+1], jstart[j] = #{b < nb : thr[b] <= j*j} - 1.
 *     Both are produced on the host from the FP64 edges (jax_powspec_amd.binning).
 *   - mode tables: int64 modes[nb] = number of stored half-spectrum entries in bin b
 *     (kx, ky in [0, n), kz in [0, n/2], each once, k = 0 included when its bin exists) and
 *     FP64 ksum[nb] = sum of |k| over them in units of k_F (needed with PK_FLAG_KMEAN only).
 *     They depend on n and thr only and are exact (jax_powspec_amd.binning.bin_modes).
 *     The per-bin counts an estimate returns are not copied from modes: the binning pass
 *     counts the entries it bins, on the same pass as the sums (src/correlations.py:48), and
 *     they normalise P_l; modes is read only for the mean |k| (ksum / count).
 */
#ifndef PK_API_H
#define PK_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PK_API_VERSION 1

#define PK_OK 0
#define PK_E_INVALID (-1)      /* bad argument (size, stride, unsupported n, ...) */
#define PK_E_OOM (-2)          /* device allocation failed (plan tables only) */
#define PK_E_LAUNCH (-3)       /* a HIP call or kernel launch failed */
#define PK_E_UNSUPPORTED (-4)  /* valid request this build does not implement */
#define PK_E_WORKSPACE (-5)    /* workspace smaller than the queried size */

/* mass-assignment stencils */
#define PK_STENCIL_CIC_REF 0   /* reference stencil of src/mas.py:142-151 (incl. the :149 corner) */
#define PK_STENCIL_CIC 1       /* textbook mass-conserving periodic CIC */
#define PK_STENCIL_TSC 2       /* textbook periodic TSC */

/* +1-neighbour wrap rules of the reference stencil (ignored by CIC/TSC, which are periodic) */
#define PK_WRAP_VEC_TRUE 0     /* src/mas.py:119-122 (cic_mas_vec, wrap=True) */
#define PK_WRAP_VEC_FALSE 1    /* src/mas.py:123-134 (cic_mas_vec, wrap=False) */
#define PK_WRAP_SCAN_TRUE 2    /* src/mas.py:20-28   (cic_mas, wrap=True) */
#define PK_WRAP_SCAN_FALSE 3   /* src/mas.py:30-38   (cic_mas, wrap=False) */

/* flags for pk_powspec_mesh / pk_estimate */
#define PK_FLAG_ZERO_DC 1      /* treat delta_k at k = 0 as exactly 0 (contrast was applied) */
#define PK_FLAG_KMEAN 2        /* also accumulate sum sqrt(m) per bin (mode-averaged k) */

/* stats buffer (device double[4]) written by pk_scatter / pk_estimate:
 *   [0] deposited mass M, [1] n^3 / M (1 / mean density), [2] mass added by the global-
 *   atomic path (scratch), [3] sum(w) / n^3 of the sorted path (scratch) */
#define PK_STATS_LEN 4

typedef struct pk_plan pk_plan;

int pk_version(void);
const char* pk_last_error(void);

/* --- plans ------------------------------------------------------------------------- */
/* Supported n: even, 8 <= n <= 1024, n = 2^a or 3 * 2^a. */
int pk_plan_create(int n, pk_plan** plan);
int pk_plan_destroy(pk_plan* plan);
int pk_supported_n(int n);   /* 1 if n is supported, else 0 */

/* --- mass assignment ----------------------------------------------------------------
 * Replaces cic_mas_vec (src/mas.py:88-153) and cic_mas (src/mas.py:5-87); CIC/TSC are
 * extensions.  mesh_out = mesh_in (or zeros if mesh_in is NULL) + deposits; mesh_out may
 * alias mesh_in.  Positions are element-strided (stride 3 reads an (N_p, 3) array in
 * place); w may be NULL (unit weights).  tile[3] = (1, ty, n) row-bundle tile of the sorted
 * LDS-privatised deposit (ty rows of full z extent; tile[2] may be 0 for n); tile == NULL
 * picks the default (ty = largest divisor of n that is <= 32 and keeps ty n 8 bytes within
 * 128 KB: 32 up to n = 512, 16 at 768 and 1024); tile[0] == 0 selects direct global
 * atomics.  Every n uses the sorted deposit unless direct is asked for: when the sort's
 * per-key tables exceed LDS they live in HBM.  stats (nullable) receives the deposited mass (PK_STATS_LEN doubles).
 * Index/fraction arithmetic follows the reference in float32: inv = 1 / ((float)box /
 * (float)n), pos = (x - xmin) * inv.
 */
int pk_scatter_workspace_size(int n, int64_t np, int stencil, const int* tile, size_t* bytes);
int pk_scatter(const float* x, int64_t sx, const float* y, int64_t sy, const float* z, int64_t sz,
               const float* w, int64_t sw, int64_t np,
               float xmin, float ymin, float zmin, double box_size, int n,
               int stencil, int wrap_mode, const int* tile,
               const float* mesh_in, float* mesh_out, double* stats,
               void* workspace, size_t workspace_bytes, void* stream);

/* --- R2C FFT ------------------------------------------------------------------------
 * Replaces jnp.fft.rfftn(delta) of src/correlations.py:22 (unnormalised, e^{-i k.x}).
 * If stats is non-NULL the contrast delta = rho * stats[1] - 1 (tests/lognormal.py:79-80)
 * is applied while the mesh is read.  Writes the spectrum layout described above.
 */
int pk_r2c(pk_plan* plan, const float* mesh, const double* stats, void* spectrum, void* stream);

/* --- P_l(k) from a mesh -------------------------------------------------------------
 * Replaces powspec_vec (src/correlations.py:7-56) and powspec_vec_fundamental
 * (src/correlations.py:60-117) together with the caller-side contrast and shot noise.
 * Runs the z/y FFT passes into `spectrum` (caller buffer, (n/2+1) n^2 float2), then the
 * x pass fused with deconvolution (x / sin x)^(2p) per axis, |delta_k|^2, Legendre
 * weights and FP64 per-bin sums.  Outputs (device, nb entries; any may be NULL except
 * pk): sums[nb][3] FP64 (S0, S2, S4), counts[nb] int64 (stored modes the pass binned, the
 * k = 0 mode included when its bin exists), pk[nb][3] float32
 * = S_l / N * (2l+1) * (box/n^2)^3 with P_0 -= shot_noise (NaN for empty bins),
 * nmodes[nb] float32, kmean[nb] FP64 (needs PK_FLAG_KMEAN; units of k_F).
 */
int pk_powspec_workspace_size(pk_plan* plan, int nb, size_t* bytes);
int pk_powspec_mesh(pk_plan* plan, const float* mesh, const double* stats,
                    const int32_t* thr, const int32_t* jstart, const int64_t* modes, const double* ksum,
                    int nb, int window_p, int flags,
                    double box_size, double shot_noise, void* spectrum,
                    double* sums, int64_t* counts, float* pk, float* nmodes, double* kmean,
                    void* workspace, size_t workspace_bytes, void* stream);

/* --- P_l(k) from a mesh without allocation ---------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * computation and outputs as pk_powspec_mesh, with the workspace replaced by a fixed
 * scratch array of PK_NOALLOC_SCRATCH_BYTES that the library holds on every device it runs
 * on until the process ends (its address is resolved when a plan is created).  The call
 * allocates nothing; the caller still passes the spectrum and output buffers.  Fails with
 * PK_E_WORKSPACE when pk_powspec_workspace_size(plan, nb) exceeds the scratch (large nb).
 * Calls that overlap on one device share the scratch and race: serialise them (one stream,
 * or events).
 */
#define PK_NOALLOC_SCRATCH_BYTES (16u << 20)
int pk_powspec_mesh_noalloc(pk_plan* plan, const float* mesh, const double* stats,
                            const int32_t* thr, const int32_t* jstart, const int64_t* modes, const double* ksum,
                            int nb, int window_p, int flags,
                            double box_size, double shot_noise, void* spectrum,
                            double* sums, int64_t* counts, float* pk, float* nmodes, double* kmean, void* stream);

/* --- whole estimate -----------------------------------------------------------------
 * particles -> assignment (mesh_in = zeros) -> contrast folded into the FFT read ->
 * P_l(k).  Workspace holds the mesh, the spectrum and all scratch.  Results are bitwise
 * reproducible: the deposit and the bin sums add integers (64-bit fixed point), so neither
 * depends on the order of atomics, and batched estimates of the same catalogue equal
 * pk_estimate's bit for bit.
 */
int pk_estimate_workspace_size(pk_plan* plan, int64_t np, int stencil, const int* tile, int nb,
                               size_t* bytes);
int pk_estimate(pk_plan* plan,
                const float* x, int64_t sx, const float* y, int64_t sy, const float* z, int64_t sz,
                const float* w, int64_t sw, int64_t np,
                float xmin, float ymin, float zmin, double box_size,
                int stencil, int wrap_mode, const int* tile,
                const int32_t* thr, const int32_t* jstart, const int64_t* modes, const double* ksum,
                int nb, int window_p, int flags, double shot_noise,
                double* sums, int64_t* counts, float* pk, float* nmodes, double* kmean,
                double* stats, void* workspace, size_t workspace_bytes, void* stream);

/* --- backward (reverse-mode derivative of an estimate) ------------------------------
 * Replaces jax.value_and_grad through cic_mas_vec + powspec_vec (tests/lognormal.py:101):
 * given dLoss/dP_l (grad_pk[nb][3], FP64) of a pk_estimate call, returns dLoss/d position
 * (grad_pos[np][3], float32, physical units) and optionally dLoss/d weight (grad_w[np]).
 * The particle, tile, bin and window arguments must be those of the forward call;
 * fwd_workspace is that call's workspace, unchanged since (it holds the spectrum after the
 * y pass), stats and counts its outputs.  The shot-noise term is constant and has no
 * derivative; the k = 0 mode is 0 in the forward and contributes none.  Through the
 * density contrast, the mean density's dependence on the particles is included.
 * No reference counterpart exists for the hand-written adjoint itself: the reference
 * differentiates its JAX code automatically.
 */
int pk_estimate_backward_workspace_size(pk_plan* plan, int64_t np, int nb, size_t* bytes);
int pk_estimate_backward(pk_plan* plan,
                         const float* x, int64_t sx, const float* y, int64_t sy, const float* z, int64_t sz,
                         const float* w, int64_t sw, int64_t np,
                         float xmin, float ymin, float zmin, double box_size,
                         int stencil, int wrap_mode, const int* tile,
                         const int32_t* thr, const int32_t* jstart, int nb, int window_p,
                         const void* fwd_workspace, size_t fwd_workspace_bytes,
                         const double* stats, const double* grad_pk, const int64_t* counts,
                         float* grad_pos, float* grad_w, void* workspace, size_t workspace_bytes, void* stream);

/* --- batches: many catalogues of one mesh side per launch --------------------------
 * No reference counterpart: the reference estimates one catalogue per call.  A batch
 * runs up to max_jobs estimates (each as pk_estimate: its own particles, stencil, window,
 * bins, shot noise and outputs) with one launch per stage for all of them: blockIdx.y
 * selects the job in a device-resident job table.  Every job's scratch (spectrum, sort
 * buffers, bin partials) is a fixed slot of the caller's workspace sized for max_np
 * particles and max_nb bins, so a batch can be refilled (pk_batch_set_jobs) and re-run
 * without allocation.  With PK_BATCH_GRAPH the launch sequence is captured into a HIP
 * graph on the first run of each job-count shape and replayed afterwards.
 *   - all jobs use the sorted + fused assignment path (tile == NULL: default rows per
 *     tile; tile = (1, ty, 0) with a compiled fused variant); pk_batch_create fails with
 *     PK_E_UNSUPPORTED when the tile has no compiled fused variant;
 *   - pk_batch_set_jobs copies the host job array into pinned memory and enqueues its
 *     upload on `stream`; the particle and output pointers must stay valid until the
 *     runs that use them are complete;
 *   - a batch is used by one stream at a time; destroy it only after its work is done.
 */
typedef struct pk_job {
  const float* x; int64_t sx;
  const float* y; int64_t sy;
  const float* z; int64_t sz;
  const float* w; int64_t sw;            /* w may be NULL (unit weights) */
  int64_t np;
  float xmin, ymin, zmin;
  double box_size;
  int stencil, wrap_mode, window_p, flags;
  const int32_t* thr; const int32_t* jstart; const int64_t* modes; const double* ksum; int nb;
  double shot_noise;
  double* sums; int64_t* counts; float* pk; float* nmodes; double* kmean;  /* as pk_estimate */
  double* stats;                          /* nullable: PK_STATS_LEN doubles */
} pk_job;

typedef struct pk_batch pk_batch;

#define PK_BATCH_GRAPH 1
/* size the sort launches for the batch capacity (max_np, TSC) instead of the current jobs:
 * fewer distinct launch shapes, so graph replays repeat when job sizes vary */
#define PK_BATCH_CAPACITY_GRIDS 2

int pk_batch_workspace_size(pk_plan* plan, int max_jobs, int64_t max_np, int max_nb, const int* tile,
                            size_t* bytes);
int pk_batch_create(pk_plan* plan, int max_jobs, int64_t max_np, int max_nb, const int* tile, int flags,
                    void* workspace, size_t workspace_bytes, pk_batch** batch);
int pk_batch_set_jobs(pk_batch* batch, int njobs, const pk_job* jobs, void* stream);
int pk_batch_run(pk_batch* batch, void* stream);
/* resolved tile (1, ty, n) of the batch's fused deposit; host int[3] */
int pk_batch_tile(pk_batch* batch, int* tile);
/* run-time tuning of a batch (the autotuner's knobs; no reference counterpart):
 * count_rounds (1..64, default 4) sets the counting chunk so that all jobs' particles take
 * about that many rounds of one block per CU (takes effect at the next pk_batch_set_jobs);
 * ypass_pct / xpass_pct (10..400, default 100) scale the y pass's and the x pass's blocks
 * per job against their resident-capacity defaults.  Drops captured graphs. */
int pk_batch_set_tuning(pk_batch* batch, int count_rounds, int ypass_pct, int xpass_pct);
int pk_batch_destroy(pk_batch* batch);

/* --- configuration-space multipoles ----------------------------------------------------
 * xi_vec(delta, box_size, s_edges) (src/correlations.py:119-187) and
 * xi_vec_fundamental(delta, box_size) (:191-260): xi(s) = irfftn(|rfftn(delta) W|^2) binned
 * by |s| over the full n^3 separation grid with Legendre weights of mu = s_z / |s|.
 *   - bins: thresholds on m = sx^2 + sy^2 + sz^2 (signed grid separations) as for P_l, with
 *     edges in grid units (s_edges n / L); modes[nb] = cells of the full grid per bin
 *     (jax_powspec_amd.binning.bin_modes(thr, n, full=True)); 1 <= nb <= PK_XI_MAX_BINS;
 *   - outputs: sums FP64 [nb][3] = sum over the bin's cells of xi_cell L_l; xi float32
 *     [nb][3] = (2l+1) sums / (N n^3); nmodes float32 [nb];
 *   - empty bins: xi 0 and nmodes +inf (xi_vec), or with PK_XI_EMPTY_NAN xi NaN and
 *     nmodes 0 (xi_vec_fundamental); the bin holding s = 0 has NaN l = 2, 4 (mu = 0/0 in
 *     the reference);
 *   - the bin sums are accumulated in 64-bit fixed point (2^62 / (2 n^3 xi(0)) units),
 *     so results are identical from run to run; workspace: pk_xi_workspace_size.
 */
#define PK_XI_MAX_BINS 1024
#define PK_XI_EMPTY_NAN 4
int pk_xi_workspace_size(pk_plan* plan, int nb, size_t* bytes);
int pk_xi_mesh(pk_plan* plan, const float* mesh, const int32_t* thr, const int32_t* jstart, const int64_t* modes,
               int nb, int window_p, int flags, double* sums, float* xi, float* nmodes, void* workspace,
               size_t workspace_bytes, void* stream);

/* --- bispectrum -------------------------------------------------------------------------
 * bispec(delta, box_size, k1, k2, theta) (src/correlations.py:334-458): for each k-shell i
 * (k1, k2, then k3(theta)), delta_i = irfftn(shell delta_k W) and I_i = irfftn(shell);
 * P_i = sum delta_i^2 / sum I_i^2 (L/n^2)^3, and for i >= 2
 * B = sum delta1 delta2 delta_i / sum I1 I2 I_i (L^2/n^3)^3, Q = B / (P1P2 + P1Pi + P2Pi).
 * The real-space sums of shells i >= 2 are evaluated in Fourier space (Parseval, against
 * the spectra of delta1 delta2 and I1 I2), which is exact up to rounding.
 *   - shell_lo / shell_hi: HOST int32[nshell], shell i = {lo_i <= m < hi_i} on
 *     m = kx^2 + ky^2 + kz^2 (jax_powspec_amd.correlations derives them from
 *     k_i / kF -/+ 1 in FP64); 2 <= nshell <= PK_BISPEC_MAX_SHELLS;
 *   - window_p: the reference multiplies delta_k by prod (x / sin x)^2, i.e. p = 1;
 *   - i_sums (nullable, device FP64 [nshell][2]): sum I_i^2 and sum I1 I2 I_i (0 for i < 2),
 *     which depend only on n and the shells, e.g. columns 1 and 3 of an earlier call's sums;
 *     when given, the indicator fields are not transformed (half the inverse FFTs);
 *   - outputs: sums FP64 [nshell][4] (sum delta^2, sum I^2, triple sums), pk float32
 *     [nshell], bk and qk float32 [nshell - 2]; workspace: pk_bispec_workspace_size.
 */
#define PK_BISPEC_MAX_SHELLS 1024
int pk_bispec_workspace_size(pk_plan* plan, int nshell, size_t* bytes);
int pk_bispec_mesh(pk_plan* plan, const float* mesh, const int32_t* shell_lo, const int32_t* shell_hi, int nshell,
                   int window_p, double box_size, const double* i_sums, double* sums, float* pk, float* bk,
                   float* qk, void* workspace, size_t workspace_bytes, void* stream);

/* --- seeded lognormal mocks -----------------------------------------------------------
 * The reference's mock pipeline (tests/create_lognormal.py:42-55): a Gaussian random
 * field from a tabulated P(k) (src/gauss_field.py:5-80), its inverse FFT
 * (scipy.fft.irfftn, :47), the lognormal transform exp(b1 g) normalised to mean 1
 * (:50-51) and Poisson population with triangular in-cell offsets
 * (src/gauss_field.py:83-110, src/populate_field.py:4-29).
 * Random numbers come from counter-based Philox4x32-10 streams keyed by the 64-bit seed
 * (the reference uses NumPy's MT19937 and JAX's threefry, so a given seed does not give
 * the reference's values): mode m = (kxx n + kyy)(n/2+1) + kzz draws from block
 * (m, 0, 1), cell c (C order) from blocks (c, j, 2), particle p from block (p, 0, 3),
 * each block with counter words (index mod 2^32, index >> 32, j, tag) and key
 * (seed mod 2^32, seed >> 32).  Results do not depend on launch shapes or devices.
 *   - kf, pkf: FP64 device arrays of nk >= 2 entries, kf strictly increasing; P(|k|) is
 *     interpolated linearly between the bracketing entries (end segments extrapolate, as
 *     the reference) and must be >= 0 wherever a mode falls (sqrt of a negative is NaN).
 *   - the spectrum is the plan's layout [n/2+1][n][n] ([kz][kx][ky]): entry (kz, i, j)
 *     equals the reference's delta_k[i, j, kz]; workspace: pk_gaussian_workspace_size.
 */
#define PK_FIELD_DELTA 0   /* the Gaussian field itself */
#define PK_FIELD_PSI_Z 1   /* its Zel'dovich displacement along z, i k_z delta / k^2 (length
                              units, k_z = n/2 plane zeroed); no reference counterpart that runs:
                              src/populate_field.py:57-69 is the broken field-level version */
int pk_gaussian_workspace_size(int n, size_t* bytes);
/* gaussian_field(grid, kf, Pkf, Rayleigh_sampling, seed, BoxSize), src/gauss_field.py:5-80 */
int pk_gaussian_field(pk_plan* plan, double box_size, const double* kf, const double* pkf, int nk, int rayleigh,
                      uint64_t seed, void* spectrum, void* workspace, size_t workspace_bytes, void* stream);
/* mesh = numpy.fft.irfftn(spectrum) (tests/create_lognormal.py:47): 1/n^3 normalisation,
 * imaginary parts of the kz = 0 and kz = n/2 entries ignored after the x and y inverse
 * transforms.  scratch is a spectrum-sized buffer and may equal spectrum (then the
 * spectrum is overwritten); otherwise the spectrum is only read. */
int pk_c2r(pk_plan* plan, const void* spectrum, void* scratch, float* mesh, void* stream);
/* irfftn of the generated field (component PK_FIELD_DELTA) or of its z displacement
 * (PK_FIELD_PSI_Z) without writing the spectrum: the modes are produced inside the first
 * inverse pass.  Equals pk_gaussian_field followed by pk_c2r (and the psi_z factor). */
int pk_gaussian_mesh(pk_plan* plan, double box_size, const double* kf, const double* pkf, int nk, int rayleigh,
                     uint64_t seed, int component, void* scratch, float* mesh, void* workspace,
                     size_t workspace_bytes, void* stream);
/* populate_field(rho, n_bins, box_size, density, seed), src/gauss_field.py:90-110: in two
 * calls on the same stream, workspace and arguments.  With bias != 0 the field is a
 * Gaussian field g and rho = exp(bias g) (tests/create_lognormal.py:50); with bias = 0 it is
 * rho itself.  Rates lambda = rho (L/n)^3 density / mean(rho) (FP64); Poisson counts per
 * cell (rates <= 0 or NaN give 0).  pk_populate_count writes the particle total to *total
 * (device int64).  pk_populate_place writes min(total, capacity) particles as float32
 * pos[p][3] in C order of their cells: cell centre + triangular offset, plus
 * growth_rate * psi_z[cell] along z when psi_z is given (redshift-space shift), wrapped
 * into [0, L).  The reference orders cells by decreasing rho instead (a permutation). */
int pk_populate_workspace_size(int n, size_t* bytes);
int pk_populate_count(const float* field, int n, double box_size, double density, double bias, uint64_t seed,
                      void* workspace, size_t workspace_bytes, int64_t* total, void* stream);
int pk_populate_place(const float* field, const float* psi_z, double growth_rate, int n, double box_size,
                      double bias, uint64_t seed, void* workspace, size_t workspace_bytes, float* pos,
                      int64_t capacity, void* stream);

/* --- per-stage timing (bench / profiling; no reference counterpart) ----------------
 * When enabled, each kernel stage is bracketed by HIP events on its own stream.
 * pk_timing_read synchronises those events, returns total ms and launch counts per stage
 * (arrays of PK_NUM_STAGES) and clears the record.  pk_timing_select(stage) restricts the
 * events to one stage (-1: every stage), so a timed loop carries two event records per
 * estimate instead of two per stage.
 */
#define PK_NUM_STAGES 15
int pk_timing_enable(int on);
int pk_timing_select(int stage);
int pk_timing_read(double* ms, int64_t* launches, int nstages);
const char* pk_stage_name(int stage);

#ifdef __cplusplus
}
#endif
#endif /* PK_API_H */
