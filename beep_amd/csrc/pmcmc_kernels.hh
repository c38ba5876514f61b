// Particle-filter kernels of the PMCMC likelihood estimate: the bootstrap
// resampling step, the state moves between sections and the gather of one
// sampled trajectory from its ancestors.  Weights and the log-mean-exp are
// FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_rng.hh"

namespace beepdev {

#define BEEP_BOOT_BLOCK 1024
#define BEEP_BOOT_UNSET 0x7FFFFFFF

// Exclusive prefix sum of one int per thread over the workgroup; returns the
// total in *total (all threads).  red holds BEEP_BOOT_BLOCK/64 ints.
__device__ inline int block_excl_scan(int v, int* red, int* total) {
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o);
    if (l >= o) x += y;
  }
  __syncthreads();
  if (l == 63) red[w] = x;
  __syncthreads();
  int before = 0, all = 0;
  for (int k = 0; k < nw; k++) {
    if (k < w) before += red[k];
    all += red[k];
  }
  *total = all;
  return before + x - v;
}

__device__ inline double block_max_d(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double t = red[0];
  for (int k = 1; k < nw; k++) t = fmax(t, red[k]);
  return t;
}

// PMCMC::bootstrap (src/pmcmc.cc:211-268) for one section, one workgroup of
// BEEP_BOOT_BLOCK threads over n particles.  L[p * Lstride] is particle p's
// section log-likelihood.
//   w_p = exp(invT (L_p - Lmax)),  wsum_p = sum_{q<=p} w_q,
//   obprob += invT Lmax + log(wsum_{n-1} / n)                 (224-236)
//   draw j: z = u_j wsum_{n-1}; pp = first p with z <= wsum_p  (245-248)
//   the first draw of each pp keeps it in place (bp[pp] = pp); later draws of
//   the same pp are "extras", handed out last-first to the particles no draw
//   chose, in increasing particle order                          (249-260)
// The linear search of the reference and the binary search here return the
// same index on the nondecreasing wsum.  Thread t owns the contiguous chunk
// [t*cs, (t+1)*cs) of particles and of draws, so the scans below keep draw
// order.  Scratch: wsum[n] double, first[n], draw[n], extra[n] int.
__global__ void __launch_bounds__(BEEP_BOOT_BLOCK) k_bootstrap(int n, const double* __restrict__ L, int Lstride,
                                                             double invT, uint64_t seed, uint32_t stream, int sec,
                                                             int* __restrict__ bp, double* __restrict__ wsum,
                                                             int* __restrict__ first, int* __restrict__ draw,
                                                             int* __restrict__ extra, double* __restrict__ obprob,
                                                             int* __restrict__ err) {
  __shared__ double redd[BEEP_BOOT_BLOCK / 64];
  __shared__ int redi[BEEP_BOOT_BLOCK / 64];
  __shared__ double s_sum;
  const int t = threadIdx.x;
  const int cs = (n + blockDim.x - 1) / blockDim.x;
  const int lo = min(n, t * cs), hi = min(n, lo + cs);

  // Lmax; a NaN anywhere makes the estimate NaN (reference: emsgEC("PMCMC",1)).
  double mx = -INFINITY;
  int bad = 0;
  for (int p = lo; p < hi; p++) {
    double v = L[(size_t)p * Lstride];
    if (v != v) bad = 1;
    mx = fmax(mx, v);
    first[p] = BEEP_BOOT_UNSET;
  }
  const double Lmax = block_max_d(mx, redd);

  // Inclusive prefix sum of the weights: chunk sums, then a scan over chunks.
  double loc = 0;
  for (int p = lo; p < hi; p++) {
    loc += exp(invT * (L[(size_t)p * Lstride] - Lmax));
    wsum[p] = loc;
  }
  __shared__ double chunk[BEEP_BOOT_BLOCK];
  chunk[t] = loc;
  __syncthreads();
  if (t < 64) {                      // wavefront 0 scans the chunk sums
    const int per = blockDim.x >> 6;
    double a = 0;
    for (int k = 0; k < per; k++) a += chunk[t * per + k];
    double x = a;
    for (int o = 1; o < 64; o <<= 1) {
      double y = __shfl_up(x, o);
      if (t >= o) x += y;
    }
    double base = x - a;
    for (int k = 0; k < per; k++) {
      double v = chunk[t * per + k];
      chunk[t * per + k] = base;
      base += v;
    }
  }
  __syncthreads();
  const double off = chunk[t];
  for (int p = lo; p < hi; p++) wsum[p] += off;
  __syncthreads();
  if (t == 0) s_sum = wsum[n - 1];
  __syncthreads();
  const double sum = s_sum;         // z = u * sum <= wsum[n-1] for u < 1
  if (t == 0) {
    double ob = invT * Lmax + log(sum / n);
    if (bad || !(ob == ob) || isinf(ob)) atomicOr(err, 4);
    else obprob[0] += ob;
  }
  __syncthreads();

  // Draws; first[pp] = the earliest draw that chose pp.
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  for (int j = lo; j < hi; j++) {
    u4 r = philox4x32_10(u4{0xFFFEu | (stream << 16), (uint32_t)j, (uint32_t)sec, 0x7FFFFFFEu}, k0, k1);
    double z = u01d(r.x) * sum;
    int a = 0, b = n;                // first index with z <= wsum[index], n if none
    while (a < b) {
      int m = (a + b) >> 1;
      if (z <= wsum[m]) b = m;
      else a = m + 1;
    }
    if (a == n) {                    // reference: emsgEC("PMCMC",2)
      atomicOr(err, 4);
      a = n - 1;
    }
    draw[j] = a;
    atomicMin(&first[a], j);
  }
  __syncthreads();

  // Extras in draw order.
  int ne = 0;
  for (int j = lo; j < hi; j++) ne += first[draw[j]] != j;
  int tot_e;
  int pos = block_excl_scan(ne, redi, &tot_e);
  for (int j = lo; j < hi; j++)
    if (first[draw[j]] != j) extra[pos++] = draw[j];
  __syncthreads();

  // Unchosen particles, in increasing order, take extras from the back.
  int nu = 0;
  for (int p = lo; p < hi; p++) nu += first[p] == BEEP_BOOT_UNSET;
  int tot_u;
  int r = block_excl_scan(nu, redi, &tot_u);
  for (int p = lo; p < hi; p++) {
    if (first[p] != BEEP_BOOT_UNSET) {
      bp[p] = p;
    } else {
      int k = tot_e - 1 - r;
      r++;
      bp[p] = k >= 0 ? extra[k] : p;
    }
  }
  if (t == 0 && tot_e != tot_u) atomicOr(err, 4);   // reference: emsgEC("PMCMC",3)
}

// Mpi::end_sec_swap (src/mpi.cc:215-327) on one device: particle p takes the
// current populations of particle bp[p].  Grid (n, chunks).  Sources are particles that kept
// their own state (bp[src] == src), so no source is overwritten.
__global__ void k_move_states(const int* __restrict__ inst, const int* __restrict__ bp, int* __restrict__ state,
                              size_t state_sz) {
  const int p = blockIdx.x;
  const int src = bp[p];
  if (src == p) return;
  const int* s_ = state + (size_t)inst[src] * state_sz;
  int* d_ = state + (size_t)inst[p] * state_sz;
  for (size_t k = blockIdx.y * blockDim.x + threadIdx.x; k < state_sz; k += (size_t)gridDim.y * blockDim.x)
    d_[k] = s_[k];
}

// Mpi::particle_sample (src/mpi.cc:331-385): the trajectory of the particle
// drawn first by the last bootstrap, traced back through the ancestry, is
// gathered section by section into instance dst.  bp holds nsec+1 rows of n
// (row k+1 written by the bootstrap of section k).  One workgroup row per
// section; the per-step trajectory slabs are tn/tm [T*AG], pt [C*AG].
__global__ void k_particle_sample(int n, int nsec, const int* __restrict__ sec_ti, const int* __restrict__ sec_tf,
                                  const int* __restrict__ bp, const int* __restrict__ inst, int dst,
                                  int* __restrict__ tn, float* __restrict__ tm, int* __restrict__ pt, size_t traj_sz,
                                  size_t pop_traj_sz, int TAG, int CAG) {
  const int sec = blockIdx.y;
  int p = bp[(size_t)nsec * n];
  for (int s = nsec - 1; s > sec; s--) p = bp[(size_t)s * n + p];
  const int src = inst[p];
  if (src == dst) return;
  const size_t t0 = sec_ti[sec], t1 = sec_tf[sec];
  const size_t a = t0 * TAG, b = t1 * TAG;
  for (size_t k = a + blockIdx.x * blockDim.x + threadIdx.x; k < b; k += (size_t)gridDim.x * blockDim.x) {
    if (tn) tn[(size_t)dst * traj_sz + k] = tn[(size_t)src * traj_sz + k];
    if (tm) tm[(size_t)dst * traj_sz + k] = tm[(size_t)src * traj_sz + k];
  }
  if (pt) {
    const size_t pa = t0 * CAG, pb = t1 * CAG;
    for (size_t k = pa + blockIdx.x * blockDim.x + threadIdx.x; k < pb; k += (size_t)gridDim.x * blockDim.x)
      pt[(size_t)dst * pop_traj_sz + k] = pt[(size_t)src * pop_traj_sz + k];
  }
}

}  // namespace beepdev

namespace beepdev {

#define BEEP_ABC_BLOCK 128
#define BEEP_ABC_TILE 64

// ABCSMC::calculate_particle_weight (src/abcsmc.cc:260-271) for n_new
// proposals at once, FP64:
//   den[i] = sum_j w[j] exp(MVN::probability(theta_i, prev_j)),
//   MVN::probability(a, b) = -0.5 sum_{v1} sum_{v2} ((a-b)_v1 / size) invM[v1][v2] ((a-b)_v2 / size)
//                                                         (src/mvn.cc:185-195)
//   out[i] = exp(logprior[i]) / den[i].
// One thread per proposal; the previous generation passes through LDS in
// tiles.  LDS: invM [nvar][nvar], this block's scaled differences e[v][t]
// = (theta_i - prev_j)_v / size, and a tile of prev and w.
__global__ void __launch_bounds__(BEEP_ABC_BLOCK) k_abcsmc_weights(int nvar, int n_prev, const double* __restrict__ prev,
                                                                 const double* __restrict__ w, int n_new,
                                                                 const double* __restrict__ theta,
                                                                 const double* __restrict__ invM, double size,
                                                                 const double* __restrict__ logprior,
                                                                 double* __restrict__ out) {
  extern __shared__ double sh[];
  double* sM = sh;                                   // [nvar][nvar]
  double* se = sM + nvar * nvar;                     // [nvar][BEEP_ABC_BLOCK]
  double* sp = se + nvar * BEEP_ABC_BLOCK;           // [BEEP_ABC_TILE][nvar]
  double* sw = sp + BEEP_ABC_TILE * nvar;            // [BEEP_ABC_TILE]
  const int t = threadIdx.x;
  const int i = blockIdx.x * BEEP_ABC_BLOCK + t;
  for (int k = t; k < nvar * nvar; k += BEEP_ABC_BLOCK) sM[k] = invM[k];
  double den = 0.0;
  for (int j0 = 0; j0 < n_prev; j0 += BEEP_ABC_TILE) {
    const int nj = min(BEEP_ABC_TILE, n_prev - j0);
    __syncthreads();
    for (int k = t; k < nj * nvar; k += BEEP_ABC_BLOCK) sp[k] = prev[(size_t)j0 * nvar + k];
    for (int k = t; k < nj; k += BEEP_ABC_BLOCK) sw[k] = w[j0 + k];
    __syncthreads();
    if (i < n_new) {
      for (int jj = 0; jj < nj; jj++) {
        for (int v = 0; v < nvar; v++) se[v * BEEP_ABC_BLOCK + t] = (theta[(size_t)i * nvar + v] - sp[jj * nvar + v]) / size;
        double sum = 0.0;
        for (int v1 = 0; v1 < nvar; v1++) {
          const double e1 = se[v1 * BEEP_ABC_BLOCK + t];
          for (int v2 = 0; v2 < nvar; v2++) sum += e1 * sM[v1 * nvar + v2] * se[v2 * BEEP_ABC_BLOCK + t];
        }
        den += sw[jj] * exp(-0.5 * sum);
      }
    }
  }
  if (i < n_new) out[i] = exp(logprior[i]) / den;
}

// MC3::swap_states (src/mc3.cc:153-215) for the chains of one launch, each
// group of chains sharing a run swapped among themselves: five sweeps over
// neighbouring pairs (i, i+1) with the same run, accepting with probability
//   al = exp(0.5 (invT_i - invT_{i+1}) (EF_i - EF_{i+1}))
// and swapping the chain labels (order) and EF.  One thread per launch, as
// the sweeps are sequential; draw k uses the Philox counter
// {0xFFFD | stream << 16, k, 0, 0x7FFFFFFD}.
// The five sweeps of neighbour swaps over one ladder of n chains (EF already
// gathered); the uniforms are Philox blocks (0xFFFD | stream << 16, draw,
// ladder, 0x7FFFFFFD), draw counting the ladder's swap trials.
__device__ inline void mc3_sweeps(int n, const int* __restrict__ run, const double* __restrict__ invT,
                                  double* __restrict__ EF, int* __restrict__ order, int* __restrict__ ntr,
                                  int* __restrict__ nac, uint64_t seed, uint32_t stream, uint32_t ladder) {
  for (int i = 0; i < n; i++) {
    order[i] = i;
    ntr[i] = 0;
    nac[i] = 0;
  }
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  uint32_t draw = 0;
  for (int loop = 0; loop < 5; loop++) {
    for (int i = 0; i + 1 < n; i++) {
      if (run[i] != run[i + 1]) continue;
      const double al = exp(0.5 * (invT[i] - invT[i + 1]) * (EF[i] - EF[i + 1]));
      ntr[i]++;
      u4 r = philox4x32_10(u4{0xFFFDu | (stream << 16), draw++, ladder, 0x7FFFFFFDu}, k0, k1);
      if (u01d(r.x) < al) {
        nac[i]++;
        int to = order[i];
        order[i] = order[i + 1];
        order[i + 1] = to;
        double te = EF[i];
        EF[i] = EF[i + 1];
        EF[i + 1] = te;
      }
    }
  }
}

__global__ void k_mc3_swap(int n, const int* __restrict__ run, const double* __restrict__ invT,
                           const double* __restrict__ EFall, const int* __restrict__ inst, double* __restrict__ EF,
                           int* __restrict__ order, int* __restrict__ ntr, int* __restrict__ nac, uint64_t seed,
                           uint32_t stream) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int i = 0; i < n; i++) EF[i] = EFall[inst[i]];
  mc3_sweeps(n, run, invT, EF, order, ntr, nac, seed, stream, 0u);
}

// Many ladders at once, one thread each: ladder l holds the chains
// ptr[l] .. ptr[l+1]-1, whose EF are read through per-chain pointers (their
// contexts' EF arrays); order is local to the ladder.
__global__ void k_mc3_swap_ladders(int nladder, const int* __restrict__ ptr, const int* __restrict__ run,
                                   const double* __restrict__ invT, const double* const* __restrict__ efp,
                                   double* __restrict__ EF, int* __restrict__ order, int* __restrict__ ntr,
                                   int* __restrict__ nac, uint64_t seed, uint32_t stream) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nladder) return;
  const int b = ptr[l], e = ptr[l + 1];
  for (int i = b; i < e; i++) EF[i] = *efp[i];
  mc3_sweeps(e - b, run + b, invT + b, EF + b, order + b, ntr + b, nac + b, seed, stream, (uint32_t)l);
}

}  // namespace beepdev
