// Counter-based random numbers and the Poisson / binomial samplers used by
// the tau-leap kernels.
//
// Streams.  Every draw is a pure function of
//   (seed, stream, instance key, step, cell = area*G + dp, transition, draw#)
// so a trajectory is the same whichever batch, slot or launch shape computes
// it.  Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
// as easy as 1, 2, 3", SC'11) maps a 128-bit counter and 64-bit key to 128
// random bits:
//   key     = seed
//   counter = (step | stream << 16, cell, word2, instance key)
// word2 = transition group q (transitions 4q..4q+3 take their first uniform
// from lanes 0..3 of that block) or, for further uniforms of transition tr,
// 0x80000000 | tr << 22 | block index.
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
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace beepdev {

struct u4 { uint32_t x, y, z, w; };

__device__ __host__ inline uint32_t mulhi32(uint32_t a, uint32_t b) {
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

__device__ __host__ inline u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int i = 0; i < 10; i++) {
    uint32_t hi0 = mulhi32(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    uint32_t hi1 = mulhi32(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    u4 o;
    o.x = hi1 ^ c.y ^ k0;
    o.y = lo1;
    o.z = hi0 ^ c.w ^ k1;
    o.w = lo0;
    c = o;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

__device__ inline uint32_t pick(const u4& b, int i) {
  return i == 0 ? b.x : (i == 1 ? b.y : (i == 2 ? b.z : b.w));
}

// Uniform on the open interval (0,1) from 32 random bits, FP64.
__device__ inline double u01d(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }
// Uniform on (0,1) from the top 24 bits, FP32.
__device__ inline float u01f(uint32_t x) { return ((float)(x >> 8) + 0.5f) * 5.9604644775390625e-08f; }

// Further uniforms of one (step, cell, transition) draw.
struct ExtraUniforms {
  uint32_t k0, k1, c0, c1, c3, tr;
  uint32_t block;
  int idx;
  u4 buf;
  __device__ inline double next() {
    if (idx >= 4) {
      u4 c = {c0, c1, 0x80000000u | (tr << 22) | block, c3};
      buf = philox4x32_10(c, k0, k1);
      block++;
      idx = 0;
    }
    uint32_t x = pick(buf, idx);
    idx++;
    return u01d(x);
  }
};

#define BEEP_MAX_EXTRA_BLOCKS 64u

// Poisson(lam) with first uniform u (FP32 path for small means).
__device__ inline int poisson_inversion(float lam, float u) {
  float p = expf(-lam);
  if (u <= p) return 0;
  float F = p;
  int k = 0;
  // The CDF is summed until it passes u.  When FP32 rounding leaves u above
  // the largest representable CDF value, the search stops where further terms
  // no longer change the sum (a tail event of probability ~1e-7).
  while (k < 64) {
    k++;
    p *= lam / (float)k;
    float Fn = F + p;
    if (u <= Fn || Fn == F) return k;
    F = Fn;
  }
  return k;
}

__device__ inline int poisson_ptrs(double lam, double ufirst, ExtraUniforms& r) {
  const double slam = sqrt(lam), loglam = log(lam);
  const double b = 0.931 + 2.53 * slam;
  const double a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  double Uraw = ufirst;
  for (;;) {
    double U = Uraw - 0.5;
    double V = r.next();
    double us = 0.5 - fabs(U);
    double k = floor((2.0 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return (int)k;
    if (!(k < 0.0 || (us < 0.013 && V > us))) {
      if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0)) return (int)k;
    }
    if (r.block >= BEEP_MAX_EXTRA_BLOCKS) return (int)floor(lam);   // unreachable in practice
    Uraw = r.next();
  }
}

__device__ inline int poisson_draw(float lam, float ufirst_f, double ufirst_d, ExtraUniforms& r) {
  if (lam < 10.0f) return poisson_inversion(lam, ufirst_f);
  return poisson_ptrs((double)lam, ufirst_d, r);
}

// log(k!) - [(k+1/2) log(k+1) - (k+1) + log(2 pi)/2]: the Stirling-series
// remainder used by BTRD's final acceptance test.
__device__ inline double stirling_fc(double k) {
  return lgamma(k + 1.0) - ((k + 0.5) * log(k + 1.0) - (k + 1.0) + 0.91893853320467274178);
}

__device__ inline int binomial_inversion(int n, double p, double ufirst, ExtraUniforms& r) {
  const double q = 1.0 - p, s = p / q, a = (n + 1) * s;
  const double r0 = pow(q, (double)n);
  double u = ufirst;
  for (;;) {
    double rr = r0;
    int k = 0;
    while (u > rr) {
      u -= rr;
      k++;
      if (k > n) break;
      rr *= (a / k - s);
      if (rr <= 0.0) { k = n + 1; break; }
    }
    if (k <= n) return k;
    if (r.block >= BEEP_MAX_EXTRA_BLOCKS) return (int)floor(n * p);   // unreachable in practice
    u = r.next();
  }
}

__device__ inline int binomial_btrd(int n, double p, double ufirst, ExtraUniforms& r) {
  const double q = 1.0 - p;
  const double m = floor((n + 1) * p);
  const double rr = p / q, nr = (n + 1) * rr, npq = n * p * q;
  const double snpq = sqrt(npq);
  const double b = 1.15 + 2.53 * snpq;
  const double a = -0.0873 + 0.0248 * b + 0.01 * p;
  const double c = n * p + 0.5;
  const double alpha = (2.83 + 5.1 / b) * snpq;
  const double vr = 0.92 - 4.2 / b;
  const double urvr = 0.86 * vr;
  double v = ufirst;
  for (;;) {
    double u;
    if (v <= urvr) {
      u = v / vr - 0.43;
      return (int)floor((2.0 * a / (0.5 - fabs(u)) + b) * u + c);
    }
    if (v >= vr) {
      u = r.next() - 0.5;
    } else {
      u = v / vr - 0.93;
      u = (u < 0 ? -0.5 : 0.5) - u;
      v = r.next() * vr;
    }
    double us = 0.5 - fabs(u);
    double k = floor((2.0 * a / us + b) * u + c);
    bool accept = false;
    if (k >= 0.0 && k <= (double)n) {
      v = v * alpha / (a / (us * us) + b);
      double km = fabs(k - m);
      if (km <= 15.0) {
        double f = 1.0;
        if (m < k) {
          for (double i = m + 1; i <= k; i += 1.0) f *= (nr / i - rr);
        } else if (m > k) {
          for (double i = k + 1; i <= m; i += 1.0) v *= (nr / i - rr);
        }
        accept = (v <= f);
      } else {
        double lv = log(v);
        double rho = (km / npq) * (((km / 3.0 + 0.625) * km + 1.0 / 6.0) / npq + 0.5);
        double t = -km * km / (2.0 * npq);
        if (lv < t - rho) accept = true;
        else if (lv <= t + rho) {
          double nm = n - m + 1;
          double h = (m + 0.5) * log((m + 1) / (rr * nm)) + stirling_fc(m) + stirling_fc(n - m);
          double nk = n - k + 1;
          accept = lv <= h + (n + 1) * log(nm / nk) + (k + 0.5) * log(nk * rr / (k + 1)) - stirling_fc(k) -
                             stirling_fc(n - k);
        }
      }
    }
    if (accept) return (int)k;
    if (r.block >= BEEP_MAX_EXTRA_BLOCKS) return (int)floor(n * p);   // unreachable in practice
    v = r.next();
  }
}

// Binomial(n, p) for 0 < p < 1.
__device__ inline int binomial_draw(int n, double p, double ufirst, ExtraUniforms& r) {
  if (n <= 0) return 0;
  bool flip = p > 0.5;
  double pp = flip ? 1.0 - p : p;
  int k;
  if (n * pp < 10.0) k = binomial_inversion(n, pp, ufirst, r);
  else k = binomial_btrd(n, pp, ufirst, r);
  return flip ? n - k : k;
}

}  // namespace beepdev
