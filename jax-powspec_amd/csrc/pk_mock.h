// Device arithmetic of the seeded lognormal mock generator: the counter-based random
// streams, the Gaussian half-spectrum mode, Poisson sampling and in-cell displacements.
//
// Every random number is a pure function of (seed, index, draw, tag) through
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), so any thread can produce any mode,
// cell or particle in any order, and a result does not depend on the launch shape.
//   field mode m = (kxx n + kyy)(n/2 + 1) + kzz : block (m, 0, TAG_FIELD)
//   Poisson cell c (C order), inversion uniform  : block (c >> 1, 0, TAG_POISSON), words
//                                                  (x, y) for even c, (z, w) for odd c
//   Poisson cell c, rejection rounds (lam >= 10) : blocks (c, j, TAG_PTRS), j = 0, 1, ...
//   particle p (output order)                    : block (p, 0, TAG_DISP)
// with key (seed mod 2^32, seed >> 32).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pk_internal.h"

namespace pk {

constexpr uint32_t TAG_FIELD = 1;
constexpr uint32_t TAG_POISSON = 2;
constexpr uint32_t TAG_DISP = 3;
constexpr uint32_t TAG_PTRS = 4;
constexpr int POISSON_INV_KMAX = 100;   // inversion bound (lambda < 10: tail beyond < 1e-60)
constexpr int POISSON_PTRS_MAX_IT = 64; // rejection rounds before the fallback floor(lambda)
constexpr double POISSON_LAM_MAX = 1073741824.0;  // rates are clamped to 2^30 (counts fit int32)

struct Philox4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                                 uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    // one 32 x 32 -> 64-bit product each (a single v_mad_u64_u32 instead of mul_hi + mul_lo)
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  return Philox4{c0, c1, c2, c3};
}

__device__ __forceinline__ Philox4 rng_block(uint64_t index, uint32_t j, uint32_t tag, uint32_t k0, uint32_t k1) {
  return philox4x32_10((uint32_t)index, (uint32_t)(index >> 32), j, tag, k0, k1);
}

// ((w >> 9) + 1/2) / 2^23: float32-exact, in (0, 1)
__device__ __forceinline__ float uniform_f32(uint32_t w) { return ((float)(w >> 9) + 0.5f) * 0x1p-23f; }
// ((a >> 6) 2^26 + (b >> 6) + 1/2) / 2^52: FP64-exact, in (0, 1)
__device__ __forceinline__ double uniform_f64(uint32_t a, uint32_t b) {
  return ((double)(a >> 6) * 67108864.0 + (double)(b >> 6) + 0.5) * 0x1p-52;
}

// Mode (kxx, kyy, kzz) of the reference's half spectrum delta_k (src/gauss_field.py:24-78):
// amplitude amp[m] (= sqrt(P (n^2/L)^3) at |k| = k_F sqrt(m)) times sqrt(-ln u) with
// Rayleigh sampling, phase 2 pi u'.  On the kz = 0 and kz = n/2 planes the first of
// (k, -k) in the reference's loop order keeps its draws and the other is its conjugate
// (:61-73); self-conjugate modes are real and equal to the amplitude (:74-75); k = 0 is 0
// (:78).
__device__ __forceinline__ float2 gauss_mode(const FieldGen& g, int kxx, int kyy, int kzz) {
  const int n = g.n, h = n >> 1;
  const int mx = kxx ? n - kxx : 0, my = kyy ? n - kyy : 0;
  const int a = kxx * n + kyy, b = mx * n + my;
  const bool plane = (kzz == 0) | (kzz == h);
  const bool conj = plane && b < a;
  const bool selfc = plane && a == b;
  const int own = conj ? b : a;
  const uint64_t m = (uint64_t)own * (uint64_t)(h + 1) + (uint64_t)kzz;
  const Philox4 w = rng_block(m, 0, TAG_FIELD, g.k0, g.k1);
  const int sx = kxx > h ? kxx - n : kxx, sy = kyy > h ? kyy - n : kyy;
  float A = g.amp[sx * sx + sy * sy + kzz * kzz];
  if (g.rayleigh) A *= sqrtf(-logf(uniform_f32(w.y)));
  if (selfc) return make_float2((a | kzz) == 0 ? 0.0f : A, 0.0f);
  float s, c;
  sincospif(2.0f * uniform_f32(w.x), &s, &c);
  return make_float2(A * c, conj ? -A * s : A * s);
}

// Zel'dovich displacement along the third axis from the same mode:
// psi_z(k) = i k_z delta(k) / k^2 (physical units; l2pi = L / 2 pi); kz = n/2 and k = 0 are 0
__device__ __forceinline__ float2 psi_z_mode(const FieldGen& g, int kxx, int kyy, int kzz) {
  const int n = g.n, h = n >> 1;
  if (kzz == h) return make_float2(0.0f, 0.0f);
  const int sx = kxx > h ? kxx - n : kxx, sy = kyy > h ? kyy - n : kyy;
  const int m = sx * sx + sy * sy + kzz * kzz;
  if (m == 0) return make_float2(0.0f, 0.0f);
  const float2 v = gauss_mode(g, kxx, kyy, kzz);
  const float f = (float)kzz * g.l2pi / (float)m;
  return make_float2(-f * v.y, f * v.x);
}

// log(k!) in FP64: exact table below 16 (the correctly rounded values), Stirling's series
// with four correction terms above (relative error < 1e-15 at k = 16)
__device__ __forceinline__ double log_factorial(double k) {
  constexpr double tab[16] = {
    0x0.0p+0,
    0x0.0p+0,
    0x1.62e42fefa39ecp-1,
    0x1.cab0bfa2a2004p+0,
    0x1.96ca77c922cf7p+1,
    0x1.326643c4479cap+2,
    0x1.a51273acf01cbp+2,
    0x1.10ce1f32dcc30p+3,
    0x1.5358e82fcb70cp+3,
    0x1.99a8921a7f7cep+3,
    0x1.e357590954d14p+3,
    0x1.180973f3a8d74p+4,
    0x1.3fcba16d50143p+4,
    0x1.68d5a9c3b32cdp+4,
    0x1.930f3df162a43p+4,
    0x1.be636a63fd346p+4};
  if (k < 16.0) return tab[(int)k];
  const double x = k + 1.0, r = 1.0 / x, r2 = r * r;
  const double corr = r * (1.0 / 12.0 - r2 * (1.0 / 360.0 - r2 * (1.0 / 1260.0 - r2 * (1.0 / 1680.0))));
  return (x - 0.5) * log(x) - x + 0.91893853320467274178 + corr;
}

// Hormann's PTRS transformed rejection (1993) for 10 <= lam <= 2^30
__device__ __forceinline__ int poisson_ptrs(double lam, uint64_t c, uint32_t k0, uint32_t k1) {
  const double slam = sqrt(lam), loglam = log(lam);
  const double b = 0.931 + 2.53 * slam;
  const double a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  const double lia = log(invalpha);
  for (int j = 0; j < POISSON_PTRS_MAX_IT; ++j) {
    const Philox4 w = rng_block(c, (uint32_t)j, TAG_PTRS, k0, k1);
    const double U = uniform_f64(w.x, w.y) - 0.5;
    const double V = uniform_f64(w.z, w.w);
    const double us = 0.5 - fabs(U);
    const double kk = floor((2.0 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return (int)kk;
    if (kk < 0.0 || (us < 0.013 && V > us)) continue;
    if (log(V) + lia - log(a / (us * us) + b) <= -lam + kk * loglam - log_factorial(kk)) return (int)kk;
  }
  return (int)floor(lam);
}

// Poisson draw of rate lam for cell c with inversion uniform u: inversion (sequential
// search of the CDF in FP64) for lam < 10, PTRS above.  lam <= 0 or NaN -> 0; lam above
// 2^30 (or infinite) is taken as 2^30.
__device__ __forceinline__ int poisson_draw(double lam, double u, uint64_t c, uint32_t k0, uint32_t k1) {
  if (!(lam > 0.0)) return 0;
  lam = fmin(lam, POISSON_LAM_MAX);
  if (lam >= 10.0) return poisson_ptrs(lam, c, k0, k1);
  double p = exp(-lam), F = p;
  int k = 0;
  while (u > F && k < POISSON_INV_KMAX) {
    ++k;
    p *= lam / (double)k;
    F += p;
  }
  return k;
}

// draws of the cell pair (c, c + 1), c even: one Philox block serves both inversions
__device__ __forceinline__ void poisson_pair(double l0, double l1, uint64_t c, uint32_t k0, uint32_t k1, int& n0,
                                             int& n1) {
  const Philox4 w = rng_block(c >> 1, 0, TAG_POISSON, k0, k1);
  n0 = poisson_draw(l0, uniform_f64(w.x, w.y), c, k0, k1);
  n1 = poisson_draw(l1, uniform_f64(w.z, w.w), c + 1, k0, k1);
}

// triangular in-cell offset of src/gauss_field.py:83-88 in cell units: R = 2u - 1,
// sign(R) (1 - sqrt|R|); R is never 0
__device__ __forceinline__ double tri_offset(uint32_t w) {
  const double R = 2.0 * (double)uniform_f32(w) - 1.0;
  return R < 0.0 ? -(1.0 - sqrt(-R)) : 1.0 - sqrt(R);
}

}  // namespace pk
