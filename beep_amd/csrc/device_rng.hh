// Counter-based random numbers and the Poisson / binomial samplers used by
// the tau-leap kernels.
//
// Streams.  Every draw is a pure function of
//   (seed, stream, instance key, step, cell = area*G + dp, transition order)
// so a trajectory is the same whichever batch, slot or launch shape computes
// it.  Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
// as easy as 1, 2, 3", SC'11) maps a 128-bit counter and 64-bit key to 128
// random bits:
//   key     = seed
//   counter = (step | stream << 16, cell, block, instance key)
// The uniforms of one (step, cell) form a pool read in order by the draws of
// its transitions, taken in transition order; block counts the 4-word Philox
// blocks the pool has used.
//
// Samplers (the reference draws with libstdc++ std::poisson_distribution and
// std::binomial_distribution, src/utils.cc:148-180; the device draws the same
// distributions with these published methods):
//   Poisson, mean < 10:  sequential-search inversion of the CDF;
//   Poisson, mean >= 10: PTRS, transformed rejection with squeeze
//                        (W. Hormann, "The transformed rejection method for
//                        generating Poisson random variables", Insurance:
//                        Mathematics and Economics 12 (1993) 39-45);
//   binomial, n*min(p,1-p) < 10: inversion (BINV, Kachitvichyanukul &
//                        Schmeiser, CACM 31 (1988) 216-222);
//   binomial otherwise:  BTRD (W. Hormann, "The generation of binomial random
//                        variates", J. Statist. Comput. Simul. 46 (1993) 101-110).
// All arithmetic is FP32: the rejection tests use C. Loader's saddle-point
// log-pmf, candidates are formed as floor(mean) + floor(frac(mean) + x) so they
// are exact integers, means of 2^23 and above are drawn as sums of Poisson
// variates (superposition) and binomials of 2^23 trials and above as sums of
// binomials over chunks of the trials, both exact in distribution.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#else
// Run-time compilation provides the fixed-width types in its own namespace.
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::uint16_t uint16_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
#endif

namespace beepdev {

struct u4 { uint32_t x, y, z, w; };

__device__ __host__ inline uint32_t mulhi32(uint32_t a, uint32_t b) {
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

__device__ __host__ inline u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int i = 0; i < 10; i++) {
    // One 32x32->64 product gives both halves (v_mad_u64_u32).
    uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    u4 o;
    o.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    o.y = (uint32_t)p1;
    o.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    o.w = (uint32_t)p0;
    c = o;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Uniform on the open interval (0,1) from 32 random bits, FP64.
__device__ inline double u01d(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }
// Uniform on (0,1) from the top 24 bits, FP32.
__device__ inline float u01f(uint32_t x) { return ((float)(x >> 8) + 0.5f) * 5.9604644775390625e-08f; }

// The uniform pool of one (step, cell).
struct CellRng {
  uint32_t k0, k1, c0, c1, c3;
  uint32_t block;
  int idx;
  u4 buf;
  __device__ inline void init(uint32_t kk0, uint32_t kk1, uint32_t cc0, uint32_t cc1, uint32_t cc3) {
    k0 = kk0; k1 = kk1; c0 = cc0; c1 = cc1; c3 = cc3; block = 0; idx = 4;
  }
  __device__ inline uint32_t bits() {
    if (idx >= 4) {
      buf = philox4x32_10(u4{c0, c1, block, c3}, k0, k1);
      block++;
      idx = 0;
    }
    // The block is consumed from .x and shifted down: no per-lane index into
    // it, which the compiler would otherwise lower to a stack array.
    uint32_t x = buf.x;
    buf.x = buf.y; buf.y = buf.z; buf.z = buf.w;
    idx++;
    return x;
  }
  __device__ inline float next() { return u01f(bits()); }
  // Fills the first block now (the same block the first bits() would make):
  // where the lanes of a wavefront make their first draw at different places
  // of an unrolled loop, this runs the generator once, converged.
  __device__ inline void prime() {
    buf = philox4x32_10(u4{c0, c1, block, c3}, k0, k1);
    block++;
    idx = 0;
  }
};
typedef CellRng ExtraUniforms;

#define BEEP_MAX_EXTRA_BLOCKS 4096u

// Blocks a pool has used: the low 22 bits of its block word.  A cell-step's
// pool counts from 0; a large draw's own stream from 0x80000000 | tr << 22
// (draw_item_body), so the transition bits above must not count.  The
// rejection loops stop at BEEP_MAX_EXTRA_BLOCKS blocks (a bound on the loop,
// reached with probability far below 1e-100).
__device__ inline uint32_t blocks_used(const CellRng& r) { return r.block & 0x3FFFFFu; }

// ---- log-pmf pieces in FP32 (C. Loader, "Fast and accurate computation of
// binomial probabilities", 2000): log n! = (n+1/2) log n - n + log(2 pi)/2 +
// stirlerr(n), and bd0(x, m) = x log(x/m) + m - x evaluated without
// cancellation, so log Pois(k; m) = -log(2 pi k)/2 - stirlerr(k) - bd0(k, m).

__device__ inline float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// stirlerr(0..16) from lgamma (entry 0 unused).
__constant__ float kStirlerr[17] = {0.0f, 8.106146680e-02f, 4.134069596e-02f, 2.767792568e-02f, 2.079067210e-02f,
                                    1.664469119e-02f, 1.387612882e-02f, 1.189670995e-02f, 1.041126526e-02f,
                                    9.255462183e-03f, 8.330563433e-03f, 7.573675488e-03f, 6.942840107e-03f,
                                    6.408994188e-03f, 5.951370113e-03f, 5.554733552e-03f, 5.207655920e-03f};

__device__ inline float stirlerr_f(float n) {
  // The table below 17, the asymptotic series above.
  if (n <= 16.f) return kStirlerr[(int)n];
  const float S0 = 1.f / 12.f, S1 = 1.f / 360.f, S2 = 1.f / 1260.f, S3 = 1.f / 1680.f;
  float in = __builtin_amdgcn_rcpf(n), inn = in * in;
  if (n > 80.f) return (S0 - (S1 - S2 * inn) * inn) * in;
  return (S0 - (S1 - (S2 - S3 * inn) * inn) * inn) * in;
}

__device__ inline float bd0_f(float x, float m) {
  float d = x - m;
  if (fabsf(d) < 0.1f * (x + m)) {
    // 2 x sum_{j>=1} v^(2j+1)/(2j+1) with v = d/(x+m): |v| < 0.1, so 8 terms
    // reach FP32 precision.
    const float inv_odd[9] = {0.f, 1.f / 3.f, 1.f / 5.f, 1.f / 7.f, 1.f / 9.f, 1.f / 11.f, 1.f / 13.f, 1.f / 15.f,
                              1.f / 17.f};
    float v = fdiv(d, x + m);
    float s = d * v;
    float ej = 2.f * x * v;
    v = v * v;
#pragma unroll
    for (int j = 1; j < 9; j++) {
      ej *= v;
      s += ej * inv_odd[j];
    }
    return s;
  }
  return x * logf(x / m) + m - x;
}

// log Pois(k; lam), k >= 1.
__device__ inline float log_pois_f(float k, float lam) {
  return -0.5f * __logf(6.2831853071795865f * k) - stirlerr_f(k) - bd0_f(k, lam);
}

// Four terms k0+1 .. k0+4 of the inversion search with 1/k as constants:
// p_k = p_{k-1} * lam / k, F_k = F_{k-1} + p_k; r = the first k whose F_k
// reaches u (or stops growing), 0 if none.
#define BEEP_INV_BLOCK(K0)                                                              \
  {                                                                                     \
    float p1 = p * (lam * (1.0f / (K0 + 1))), F1 = F + p1;                             \
    float p2 = p1 * (lam * (1.0f / (K0 + 2))), F2 = F1 + p2;                           \
    float p3 = p2 * (lam * (1.0f / (K0 + 3))), F3 = F2 + p3;                           \
    float p4 = p3 * (lam * (1.0f / (K0 + 4))), F4 = F3 + p4;                           \
    int r = (u <= F4 || F4 == F3) ? K0 + 4 : 0;                                         \
    r = (u <= F3 || F3 == F2) ? K0 + 3 : r;                                             \
    r = (u <= F2 || F2 == F1) ? K0 + 2 : r;                                             \
    r = (u <= F1 || F1 == F) ? K0 + 1 : r;                                              \
    if (r) return r;                                                                    \
    p = p4;                                                                             \
    F = F4;                                                                             \
  }

// Poisson(lam) with first uniform u (FP32 path for small means): sequential
// inversion, the CDF summed until it passes u.  When FP32 rounding leaves u
// above the largest representable CDF value, the search stops where further
// terms no longer change the sum (a tail event of probability ~1e-7).  The
// first 16 terms are evaluated four at a time with constant reciprocals (one
// branch per four terms, which matters where the lanes of a wavefront need
// different numbers of terms); later terms, rare below mean 10, one at a time.
__device__ inline int poisson_inversion(float lam, float u) {
  float p = __expf(-lam);
  if (u <= p) return 0;
  float F = p;
  BEEP_INV_BLOCK(0)
  BEEP_INV_BLOCK(4)
  BEEP_INV_BLOCK(8)
  BEEP_INV_BLOCK(12)
  int k = 16;
  while (k < 64) {
    k++;
    p *= lam * __builtin_amdgcn_rcpf((float)k);
    float Fn = F + p;
    if (u <= Fn || Fn == F) return k;
    F = Fn;
  }
  return k;
}
#undef BEEP_INV_BLOCK

// PTRS (Hormann 1993) for 10 <= lam < 2^23: the squeeze accepts most
// candidates; the exact test compares with the saddle-point log-pmf.
__device__ inline int poisson_ptrs(float lam, float ufirst, ExtraUniforms& r) {
  const float slam = __builtin_amdgcn_sqrtf(lam);
  const float b = 0.931f + 2.53f * slam;
  const float a = -0.059f + 0.02483f * b;
  const float vr = 0.9277f - fdiv(3.6224f, b - 2.f);
  const float lfl = floorf(lam), lfr = lam - lfl;
  float Uraw = ufirst;
  for (;;) {
    float U = Uraw - 0.5f;
    float V = r.next();
    float us = 0.5f - fabsf(U);
    float k = lfl + floorf((fdiv(2.f * a, us) + b) * U + (lfr + 0.43f));
    if (us >= 0.07f && V <= vr) return (int)k;
    if (!(k < 0.f || (us < 0.013f && V > us))) {
      float lhs = __logf(V * (1.1239f + fdiv(1.1328f, b - 3.4f))) - __logf(fdiv(a, us * us) + b);
      float rhs = k == 0.f ? -lam : log_pois_f(k, lam);
      if (lhs <= rhs) return (int)k;
    }
    if (blocks_used(r) >= BEEP_MAX_EXTRA_BLOCKS) return (int)lfl;   // unreachable in practice
    Uraw = r.next();
  }
}

#define BEEP_SPLIT (8388608.0f)   // 2^23

__device__ inline int poisson_draw(float lam, float ufirst, ExtraUniforms& r) {
  if (lam < 10.0f) return poisson_inversion(lam, ufirst);
  if (lam < BEEP_SPLIT) return poisson_ptrs(lam, ufirst, r);
  // Superposition: Pois(lam) = sum of m independent Pois(lam / m).
  int m = (int)ceilf(lam / BEEP_SPLIT);
  float part = lam / (float)m;
  int n = 0;
  for (int i = 0; i < m; i++) n += poisson_ptrs(part, i == 0 ? ufirst : r.next(), r);
  return n;
}

// Inversion from k = 0 (BINV), FP32, for n * p < 10 with p <= 1/2.
__device__ inline int binomial_inversion(int n, float p, float ufirst, ExtraUniforms& r) {
  const float q = 1.f - p, s = fdiv(p, q), a = (float)(n + 1) * s;
  const float r0 = __expf((float)n * log1pf(-p));
  float u = ufirst;
  for (;;) {
    float rr = r0;
    int k = 0;
    while (u > rr) {
      u -= rr;
      k++;
      if (k > n || k > 110) break;
      rr *= (a * __builtin_amdgcn_rcpf((float)k) - s);
      if (rr <= 0.f) { k = n + 1; break; }
    }
    if (k <= n && k <= 110) return k;
    if (blocks_used(r) >= BEEP_MAX_EXTRA_BLOCKS) return (int)floorf(n * p);   // unreachable in practice
    u = (float)r.next();
  }
}

// BTRD (Hormann 1993) in FP32 for n * p >= 10, p <= 1/2, n < 2^24;
// fc(k) = stirlerr(k + 1).
__device__ inline int binomial_btrd(int n, float p, float ufirst, ExtraUniforms& r) {
  const float q = 1.f - p;
  const float nf = (float)n;
  const float m = floorf((nf + 1.f) * p);
  const float rr = fdiv(p, q), nr = (nf + 1.f) * rr, npq = nf * p * q;
  const float snpq = __builtin_amdgcn_sqrtf(npq);
  const float b = 1.15f + 2.53f * snpq;
  const float a = -0.0873f + 0.0248f * b + 0.01f * p;
  const float c = nf * p + 0.5f;
  const float alpha = (2.83f + fdiv(5.1f, b)) * snpq;
  const float vr = 0.92f - fdiv(4.2f, b);
  const float urvr = 0.86f * vr;
  float v = ufirst;
  for (;;) {
    float u;
    if (v <= urvr) {
      u = fdiv(v, vr) - 0.43f;
      return (int)floorf((fdiv(2.f * a, 0.5f - fabsf(u)) + b) * u + c);
    }
    if (v >= vr) {
      u = (float)r.next() - 0.5f;
    } else {
      u = fdiv(v, vr) - 0.93f;
      u = (u < 0.f ? -0.5f : 0.5f) - u;
      v = (float)r.next() * vr;
    }
    float us = 0.5f - fabsf(u);
    float k = floorf((fdiv(2.f * a, us) + b) * u + c);
    bool accept = false;
    if (k >= 0.f && k <= nf) {
      v = v * fdiv(alpha, fdiv(a, us * us) + b);
      float km = fabsf(k - m);
      if (km <= 15.f) {
        float f = 1.f;
        if (m < k) {
          for (float i = m + 1.f; i <= k; i += 1.f) f *= (fdiv(nr, i) - rr);
        } else if (m > k) {
          for (float i = k + 1.f; i <= m; i += 1.f) v *= (fdiv(nr, i) - rr);
        }
        accept = (v <= f);
      } else {
        float lv = __logf(v);
        float inpq = __builtin_amdgcn_rcpf(npq);
        float rho = (km * inpq) * (((km * (1.f / 3.f) + 0.625f) * km + 1.f / 6.f) * inpq + 0.5f);
        float t = -km * km * 0.5f * inpq;
        if (lv < t - rho) accept = true;
        else if (lv <= t + rho) {
          float nm = nf - m + 1.f;
          float h = (m + 0.5f) * logf((m + 1.f) / (rr * nm)) + stirlerr_f(m + 1.f) + stirlerr_f(nf - m + 1.f);
          float nk = nf - k + 1.f;
          accept = lv <= h + (nf + 1.f) * logf(nm / nk) + (k + 0.5f) * logf(nk * rr / (k + 1.f)) -
                             stirlerr_f(k + 1.f) - stirlerr_f(nf - k + 1.f);
        }
      }
    }
    if (accept) return (int)k;
    if (blocks_used(r) >= BEEP_MAX_EXTRA_BLOCKS) return (int)floorf(n * p);   // unreachable in practice
    v = (float)r.next();
  }
}

__device__ inline int binomial_chunk(int n, float p, float ufirst, ExtraUniforms& r) {
  if ((float)n * p < 10.f) return binomial_inversion(n, p, ufirst, r);
  return binomial_btrd(n, p, ufirst, r);
}

// Binomial(n, p) for 0 < p < 1.
__device__ inline int binomial_draw(int n, float p, float ufirst, ExtraUniforms& r) {
  if (n <= 0) return 0;
  bool flip = p > 0.5f;
  float pp = flip ? 1.f - p : p;
  int k = 0;
  const int chunk = 1 << 23;
  int left = n;
  bool first = true;
  while (left > 0) {          // Bin(n, p) = sum of Bin(chunk, p) over chunks of the trials
    int c = left < chunk ? left : chunk;
    k += binomial_chunk(c, pp, first ? ufirst : r.next(), r);
    first = false;
    left -= c;
  }
  return flip ? n - k : k;
}

// binomial_draw where n * min(p, 1 - p) < 10 is known: every chunk then takes
// the inversion branch, so this is the same function of (n, p, uniforms)
// without the BTRD code, for call sites inlined in the transition loop.
__device__ inline int binomial_draw_small(int n, float p, float ufirst, ExtraUniforms& r) {
  if (n <= 0) return 0;
  bool flip = p > 0.5f;
  float pp = flip ? 1.f - p : p;
  int k = 0;
  const int chunk = 1 << 23;
  int left = n;
  bool first = true;
  while (left > 0) {
    int c = left < chunk ? left : chunk;
    k += binomial_inversion(c, pp, first ? ufirst : r.next(), r);
    first = false;
    left -= c;
  }
  return flip ? n - k : k;
}

}  // namespace beepdev
