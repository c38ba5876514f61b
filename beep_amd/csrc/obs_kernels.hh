// Observation log-likelihood, fixed-state Poisson log-likelihood and the
// Metropolis-Hastings acceptance kernels.  Values are accumulated in FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_rng.hh"
#include "fused_common.hh"

namespace beepdev {

struct ObsDev {
  int nobs, ndt;
  uint32_t TAG, CAG;         // T*AG and C*AG of one step
  const int* ptr;            // [nobs+1] element ranges
  const uint32_t* elem;      // bit 31 clear: transnum index s*TAG + tr*AG + cell; set: pop index
  const int* type;           // BEEP_OBS_*
  const int* dt;             // data table
  const double* y;
  const double* sd;
  const double* shape;
  const double* factor;
  const double* thresh;
  const int* orig;           // the model's index of each stored observation
  const double* lc;          // normal types: -log(2 pi sd^2) / 2
  const double* iv;          // normal types: 1 / (2 sd^2)
  const int* ospl;           // observation spline of each stored observation, -1 = none
  int nospl, S;              // observation splines; steps (a slot holds [nospl][S] values)
};

// log Gamma(x) for x > 0, FP64: Stirling's series
//   log Gamma(x) = (x - 1/2) log x - x + log(2 pi)/2 + sum_k B_2k / (2k (2k-1) x^(2k-1))
// (Bernoulli numbers B_2..B_14) for x >= 10, where the first omitted term is
// below 4e-17, and Gamma(x) = Gamma(x + n) / (x (x+1) ... (x+n-1)) below 10.
// The arguments here (k + 1, shape + k, shape) are positive; near x = 1 and 2
// the error is ~1e-15 absolute, which is what the log-pmf sums need.
__host__ __device__ inline double lgamma_pos(double x) {
  double prod = 1.0;
  while (x < 10.0) {
    prod *= x;
    x += 1.0;
  }
  const double z = 1.0 / x, z2 = z * z;
  const double series =
      z * (1.0 / 12 + z2 * (-1.0 / 360 + z2 * (1.0 / 1260 + z2 * (-1.0 / 1680 + z2 * (1.0 / 1188 +
      z2 * (-691.0 / 360360 + z2 * (1.0 / 156)))))));
  return (x - 0.5) * log(x) - x + 0.91893853320467274178 + series - log(prod);
}

__device__ inline double poisson_lpmf(double k, double lam) {          // src/utils.cc:157-170
  if (lam == 0) return k == 0 ? 0.0 : -1e9;
  if (k < 0) return -1e9;
  if (k == 0) return -lam;
  return k * log(lam) - lam - lgamma_pos(k + 1.0);
}

__device__ inline double negbin_lpmf(double k, double m, double shape) {   // src/utils.cc:124-127
  return lgamma_pos(shape + k) - lgamma_pos(k + 1.0) - lgamma_pos(shape) + shape * log(shape / (shape + m)) +
         k * log(m / (shape + m));
}

__device__ inline double normal_lpdf(double x, double mean, double var) {  // src/utils.cc:107-111
  return -0.5 * log(6.283185307179586476925 * var) - (x - mean) * (x - mean) / (2 * var);
}

// ObservationModel::obs_prob (src/obsmodel.cc:245-305).
__device__ inline double obs_prob(double value, double y, double sd, int type, double shape, double thresh) {
  if (value < 0) value = 0;
  if (y == 999999997.0) return 0.0;
  if (y == 999999998.0) {
    if (type <= 1) return value < thresh ? normal_lpdf(thresh, thresh, sd * sd) : normal_lpdf(thresh, value, sd * sd);
    double lam = value < 1 ? 1 : value, sum = 0;
    for (double k = 0; k <= thresh; k += 1.0)
      sum += exp(type == 2 ? poisson_lpmf(k, lam) : negbin_lpmf(k, lam, shape));
    return log(sum);
  }
  if (type <= 1) return normal_lpdf(y, value, sd * sd);
  double k = floor(y + 0.5), lam = value < 1 ? 1 : value;
  if (type == 2) return poisson_lpmf(k, lam);
  return negbin_lpmf(k, lam, shape);
}

// splv: the instance's slot's observation spline values [nospl][S], or null
// when the model has none.  With a spline, each transition count is weighted
// by the spline's value at its step (src/obsmodel.cc:85-88).
__device__ inline double obs_value(const ObsDev& od, int i, const int* tn, const int* pt, const double* splv) {
  double v = 0, f = od.factor[i];
  const int os = splv ? od.ospl[i] : -1;
  for (int e = od.ptr[i]; e < od.ptr[i + 1]; e++) {
    uint32_t x = od.elem[e];
    int num = (x & 0x80000000u) ? (pt ? pt[x & 0x7FFFFFFFu] : 0) : tn[x];
    if (num != 0) v += os < 0 ? f * (double)num : f * (double)num * splv[(size_t)os * od.S + x / od.TAG];
  }
  return v;
}

// Global-typed view of a pointer held in a descriptor: a plain pointer read
// from memory is generic, and loads through it would be flat loads.
template <typename T>
__device__ inline const __attribute__((address_space(1))) T* gview(const T* p) {
  return (const __attribute__((address_space(1))) T*)p;
}

__device__ inline double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double t = 0;
  int nw = (blockDim.x + 63) >> 6;
  for (int k = 0; k < nw; k++) t += red[k];
  return t;
}

// Count of element x: a transnum index, or with bit 31 set a population
// index (zero when the population trajectory is not kept).  POP false: the
// model has no population elements.
template <bool POP>
__device__ inline int obs_elem(uint32_t x, const int* tnb, const int* ptb) {
  if (!POP) return tnb[x];
  return (x & 0x80000000u) ? (ptb ? ptb[x & 0x7FFFFFFFu] : 0) : tnb[x];
}

// Instances per workgroup of the observation kernel: four, or one when the
// per-data-table sums are wanted (eight more accumulators per instance).
template <bool WANT_DT>
constexpr int OBS_INSTANCES = WANT_DT ? 1 : 4;

// Defined below, after the kernels that call it.
template <bool WANT_DT, bool GENERAL, bool POP, bool SPL, int NI>
__device__ inline void obs_terms(const ObsDev& od, const int* __restrict__ sel, int kbeg, int kend, int d0,
                                 const int* (&tnb)[NI], const int* (&ptb)[NI], const double* (&splv)[NI],
                                 double (&L)[NI], double (&acc)[8]);

// One workgroup per OBS_INSTANCES instances.  out[b*(1+ndt)] = log-likelihood,
// out[b*(1+ndt)+1+d] = -2 * (sum over data table d)  (get_EF_datatable),
// written when WANT_DT.  The counts of an observation are summed as
// integers and scaled by its factor once; normal terms use the per-
// observation constants lc and iv.  GENERAL: the model has observations
// other than plain normal ones (Poisson, negative binomial, thresholds,
// unknowns); without it the kernel holds only the normal path, which keeps
// its register use, and so its occupancy, low.  POP: some observation counts
// population elements.  SPL: some observations carry an observation spline;
// their counts are weighted per step by the values of the instance's slot,
// ospl_tab + inst_slot[instance] * nospl * S.
template <bool WANT_DT, bool GENERAL, bool POP, bool SPL>
__global__ void __launch_bounds__(256) k_obs_loglik(const ObsDev* __restrict__ odp, int n, const int* __restrict__ inst,
                                                    const int* __restrict__ tn, const int* __restrict__ pt,
                                                    size_t traj_sz, size_t pop_traj_sz, const int* __restrict__ sel,
                                                    int nsel, double* __restrict__ out, double* __restrict__ EF,
                                                    int store_EF, const double* __restrict__ ospl_tab,
                                                    const int* __restrict__ inst_slot, int d0) {
  // Instances per workgroup: they share the observation structure, so one
  // thread reads an observation's elements once and gathers the counts of
  // all NI instances with independent loads.
  constexpr int NI = OBS_INSTANCES<WANT_DT>;
  __shared__ double red[8];
  const ObsDev od = *odp;
  const int b0 = blockIdx.x * NI;
  int ins[NI];
  const int* tnb[NI];
  const int* ptb[NI];
#pragma unroll
  for (int q = 0; q < NI; q++) {
    ins[q] = inst[b0 + q < n ? b0 + q : b0];        // past the end: recomputes instance b0, not written
    tnb[q] = tn + (size_t)ins[q] * traj_sz;
    ptb[q] = pt ? pt + (size_t)ins[q] * pop_traj_sz : nullptr;
  }
  const double* splv[NI];
#pragma unroll
  for (int q = 0; q < NI; q++) splv[q] = SPL ? ospl_tab + (size_t)inst_slot[ins[q]] * od.nospl * od.S : nullptr;
  double L[NI], acc[8];
  obs_terms<WANT_DT, GENERAL, POP, SPL, NI>(od, sel, 0, nsel, d0, tnb, ptb, splv, L, acc);
#pragma unroll
  for (int q = 0; q < NI; q++) {
    const double Lt = block_sum(L[q], red);
    if (threadIdx.x == 0 && b0 + q < n && d0 == 0) {
      out[(size_t)(b0 + q) * (1 + od.ndt)] = Lt;
      if (store_EF) EF[ins[q]] = -2.0 * Lt;
    }
  }
  if (WANT_DT) {
    // Data tables d0 .. d0 + 7 in this pass (the launcher makes one pass per
    // eight tables; the total above is written by the first).
    double* o = out + (size_t)b0 * (1 + od.ndt);
    const int nd = min(8, od.ndt - d0);
    for (int d = 0; d < nd; d++) {
      double t = block_sum(acc[d], red);
      if (threadIdx.x == 0) o[1 + d0 + d] = -2.0 * t;
    }
  }
}

// ---- observation log-likelihoods in chunks of observations ----
// A group is up to OBS_GROUP instances of one context; its observations are
// cut into chunks of BEEP_OBS_CHUNK (in stored order), one workgroup each, and
// each workgroup writes one partial sum per instance.  k_obs_reduce then adds
// an instance's partials in chunk order.  Work is spread over the whole device
// however few instances a context has, and the sum of an instance depends only
// on its context's observation count, not on the call or batch it is in.
#define BEEP_OBS_CHUNK 2048
constexpr int OBS_GROUP = 4;

struct ObsGroup {
  ObsDev od;                        // the context's observation structure (a copy: read
                                    // with scalar loads, as a kernel argument would be)
  const int* tn;                    // its transnum, [instance][S][T][AG]
  const int* pt;                    // its population trajectories, or null
  const double* ospl;               // its observation-spline values per slot, or null
  const int* inst_slot;             // its instances' slots (splines), or null
  const int* inst;                  // the group's instances
  unsigned long long traj_sz, pop_traj_sz;
  int ninst;                        // 1..OBS_GROUP
  int nchunk;                       // ceil(nobs / BEEP_OBS_CHUNK), at least 1
  int wg0;                          // first workgroup of the group in its launch
  int part0;                        // first partial record of the group
};

// Partial sums: record p holds OBS_GROUP doubles, one per instance of its
// group.  Grid: one workgroup per (group, chunk), groups in wg0 order.
template <bool GENERAL, bool POP, bool SPL>
__global__ void __launch_bounds__(256) k_obs_chunks(const ObsGroup* __restrict__ groups, int ngroups,
                                                    double* __restrict__ partial) {
  constexpr int NI = OBS_GROUP;
  __shared__ double red[8];
  // The group of this workgroup: the last one whose wg0 is at most blockIdx.x.
  int lo = 0, hi = ngroups - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (groups[mid].wg0 <= (int)blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const ObsGroup& gr = groups[lo];
  const ObsDev& od = gr.od;
  const int c = (int)blockIdx.x - gr.wg0;
  const int kbeg = c * BEEP_OBS_CHUNK;
  const int kend = min(kbeg + BEEP_OBS_CHUNK, od.nobs);
  int ins[NI];
  const int* tnb[NI];
  const int* ptb[NI];
  const double* splv[NI];
  // The pointers read from the group are generic; derived from global views,
  // the trajectory gathers compile to global (not flat) loads, as they do
  // from kernel arguments.
#pragma unroll
  for (int q = 0; q < NI; q++) {
    ins[q] = gview(gr.inst)[q < gr.ninst ? q : 0];  // past the group's end: instance 0 again, not written
    tnb[q] = (const int*)(gview(gr.tn) + (size_t)ins[q] * gr.traj_sz);
    ptb[q] = gr.pt ? (const int*)(gview(gr.pt) + (size_t)ins[q] * gr.pop_traj_sz) : nullptr;
    splv[q] = SPL ? (const double*)(gview(gr.ospl) + (size_t)gview(gr.inst_slot)[ins[q]] * od.nospl * od.S) : nullptr;
  }
  double L[NI], acc[8];
  obs_terms<false, GENERAL, POP, SPL, NI>(od, nullptr, kbeg, kend, 0, tnb, ptb, splv, L, acc);
#pragma unroll
  for (int q = 0; q < NI; q++) {
    const double Lt = block_sum(L[q], red);
    if (threadIdx.x == 0 && q < gr.ninst) partial[(size_t)(gr.part0 + c) * NI + q] = Lt;
  }
}

// An instance's place among the partials, and where its EF goes (or null).
struct ObsItem {
  int part0, nchunk, q;
  double* EF;
};

// out[i] = sum over the chunks of item i's partials, in chunk order; EF too.
__global__ void __launch_bounds__(256) k_obs_reduce(const ObsItem* __restrict__ items, int n,
                                                    const double* __restrict__ partial, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const ObsItem it = items[i];
  double L = 0;
  for (int c = 0; c < it.nchunk; c++) L += partial[(size_t)(it.part0 + c) * OBS_GROUP + it.q];
  out[i] = L;
  if (it.EF) *it.EF = -2.0 * L;
}

// The log-likelihood terms of observations [kbeg, kend) of the list sel (all
// stored observations when null), thread t taking every blockDim.x-th from
// kbeg + t: L[q] for the NI instances whose transnum, population trajectories
// and spline rows are tnb[q], ptb[q], splv[q]; with WANT_DT also acc[d], the
// sums per data table.
template <bool WANT_DT, bool GENERAL, bool POP, bool SPL, int NI>
__device__ inline void obs_terms(const ObsDev& od, const int* __restrict__ sel, int kbeg, int kend, int d0,
                                 const int* (&tnb)[NI], const int* (&ptb)[NI], const double* (&splv)[NI],
                                 double (&L)[NI], double (&acc)[8]) {
  const auto optr = gview(od.ptr);
  const auto elem = gview(od.elem);
  const auto factor = gview(od.factor);
  const auto oy = gview(od.y);
  const auto otype = gview(od.type);
#pragma unroll
  for (int q = 0; q < NI; q++) L[q] = 0;
#pragma unroll
  for (int d = 0; d < 8; d++) acc[d] = 0;
  for (int k = kbeg + (int)threadIdx.x; k < kend; k += blockDim.x) {
    const int i = sel ? sel[k] : k;
    long long cnt[NI];
#pragma unroll
    for (int q = 0; q < NI; q++) cnt[q] = 0;
    int e = optr[i];
    const int e1 = optr[i + 1];
    const int os = SPL ? gview(od.ospl)[i] : -1;
    double wsum[NI];                       // SPL: sum of count x spline value
#pragma unroll
    for (int q = 0; q < NI; q++) wsum[q] = 0;
    if (SPL && os >= 0) {
      // Transition elements only; the step of element x is x / TAG.
      for (; e < e1; e++) {
        const uint32_t x = elem[e];
        const size_t k = (size_t)os * od.S + x / od.TAG;
#pragma unroll
        for (int q = 0; q < NI; q++) {
          const int num = tnb[q][x];
          if (num != 0) wsum[q] += (double)num * splv[q][k];
        }
      }
    }
    // Four elements at a time: without population elements the trajectory
    // loads are unconditional and in flight together.
    for (; e + 4 <= e1; e += 4) {
      const uint32_t x0 = elem[e], x1 = elem[e + 1], x2 = elem[e + 2], x3 = elem[e + 3];
#pragma unroll
      for (int q = 0; q < NI; q++) {
        const int n0 = obs_elem<POP>(x0, tnb[q], ptb[q]), n1 = obs_elem<POP>(x1, tnb[q], ptb[q]);
        const int n2 = obs_elem<POP>(x2, tnb[q], ptb[q]), n3 = obs_elem<POP>(x3, tnb[q], ptb[q]);
        cnt[q] += (long long)n0 + n1 + n2 + n3;
      }
    }
    for (; e < e1; e++) {
      const uint32_t x = elem[e];
#pragma unroll
      for (int q = 0; q < NI; q++) cnt[q] += obs_elem<POP>(x, tnb[q], ptb[q]);
    }
    const double f = factor[i];
    const double y = oy[i];
    const int type = otype[i];
    const bool plain = !GENERAL || (type <= 1 && y != 999999997.0 && y != 999999998.0);
#pragma unroll
    for (int q = 0; q < NI; q++) {
      const double v = SPL && os >= 0 ? f * wsum[q] : f * (double)cnt[q];
      double lp;
      if (plain) {
        const double d = y - (v < 0 ? 0.0 : v);
        lp = gview(od.lc)[i] - d * d * gview(od.iv)[i];
      } else {
        lp = obs_prob(v, y, gview(od.sd)[i], type, gview(od.shape)[i], gview(od.thresh)[i]);
      }
      L[q] += lp;
      if (WANT_DT) {
        const int dt = gview(od.dt)[i] - d0;
#pragma unroll
        for (int d = 0; d < 8; d++)
          if (d == dt) acc[d] += lp;
      }
    }
  }
}

__global__ void k_obs_values(const ObsDev* __restrict__ odp, const int* __restrict__ tn, const int* __restrict__ pt,
                             const double* __restrict__ splv, double* __restrict__ out) {
  const ObsDev od = *odp;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < od.nobs) out[od.orig[i]] = obs_value(od, i, tn, pt, splv);
}

// State::likelihood (src/state.cc:812-826): one workgroup per instance.
__global__ void __launch_bounds__(256) k_pois_loglik(const int* __restrict__ inst, const int* __restrict__ tn,
                                                     const float* __restrict__ tm, size_t traj_sz,
                                                     double* __restrict__ out) {
  __shared__ double red[8];
  const int in = inst[blockIdx.x];
  const int* n = tn + (size_t)in * traj_sz;
  const float* l = tm + (size_t)in * traj_sz;
  double L = 0;
  for (size_t e = threadIdx.x; e < traj_sz; e += blockDim.x) L += poisson_lpmf((double)n[e], (double)l[e]);
  double t = block_sum(L, red);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// Mbp::get_al + MVN::MH (src/mbp.cc:334-361, src/mvn.cc:233-249), FP64.
// One acceptance: al and the decision with the proposal key's uniform.
__device__ inline void accept_one(double EFp, double EFi, double dlogq, uint32_t key, int mode, double invT,
                                  double EFcut, uint64_t seed, uint32_t stream, double* al_out, uint8_t* acc) {
  double al = 0;
  if (mode == 0) {
    if (EFp < EFcut) al = exp(dlogq);
  } else {
    al = exp(-invT * 0.5 * (EFp - EFi) + dlogq);
  }
  if (al > 1) al = 1;
  u4 blk = philox4x32_10(u4{0xFFFFu | (stream << 16), 0xFFFFFFFFu, 0x7FFFFFFFu, key}, (uint32_t)seed,
                         (uint32_t)(seed >> 32));
  double u = u01d(blk.x);
  *al_out = al;
  *acc = u < al ? 1 : 0;
}

__global__ void k_accept(int n, const int* __restrict__ init, const int* __restrict__ prop, const double* __restrict__ EF,
                         const double* __restrict__ dlogq, const uint32_t* __restrict__ key, int mode, double invT,
                         double EFcut, uint64_t seed, uint32_t stream, double* __restrict__ al_out,
                         uint8_t* __restrict__ acc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  accept_one(EF[prop[i]], EF[init[i]], dlogq[i], key[i], mode, invT, EFcut, seed, stream, al_out + i, acc + i);
}

// The same over proposals of many contexts: the EF of each proposed and
// initial instance through a pointer into its context's EF array, and a
// temperature per proposal.
__global__ void k_accept_ptrs(int n, const double* const* __restrict__ efp, const double* const* __restrict__ efi,
                              const double* __restrict__ invT, const double* __restrict__ dlogq,
                              const uint32_t* __restrict__ key, int mode, double EFcut, uint64_t seed, uint32_t stream,
                              double* __restrict__ al_out, uint8_t* __restrict__ acc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  accept_one(*efp[i], *efi[i], dlogq[i], key[i], mode, invT[i], EFcut, seed, stream, al_out + i, acc + i);
}

// ---- prediction helpers ----

// tab[dst] = tab[src] with beta[s] for steps s >= from scaled by factor (R(t)
// multiplied after a change point: beta is proportional to R(t_s)).
__global__ void k_derive_slot(float* __restrict__ tab, size_t stride, const int* __restrict__ dst,
                              const int* __restrict__ src, const float* __restrict__ factor,
                              const int* __restrict__ from, int off_beta, int S, int nR) {
  const int i = blockIdx.x;        // grid (n, chunks)
  const float* s_ = tab + (size_t)src[i] * stride;
  float* d_ = tab + (size_t)dst[i] * stride;
  const float f = factor[i];
  // beta[S][nR]: every group's beta from step from[i] on.
  const size_t first = (size_t)off_beta + (size_t)from[i] * nR, end = (size_t)off_beta + (size_t)S * nR;
  for (size_t k = blockIdx.y * blockDim.x + threadIdx.x; k < stride; k += (size_t)gridDim.y * blockDim.x) {
    float v = s_[k];
    if (k >= first && k < end) v *= f;
    d_[k] = v;
  }
}

// Parameter slots written from a packed upload: tab[slot[i]] = packed[i]
// (n blocks of `stride` floats); grid (n, chunks): slot on x.
__global__ void k_scatter_slots(float* __restrict__ tab, size_t stride, const float* __restrict__ packed,
                                const int* __restrict__ slot) {
  const int i = blockIdx.x;
  const float* s_ = packed + (size_t)i * stride;
  float* d_ = tab + (size_t)slot[i] * stride;
  for (size_t k = blockIdx.y * blockDim.x + threadIdx.x; k < stride; k += (size_t)gridDim.y * blockDim.x) d_[k] = s_[k];
}

// Blocks of a packed upload into slots of many contexts: dst[i][k] =
// packed[src_off[i] + k] for k < len[i]; grid (n, chunks).
__global__ void k_scatter_blocks(const float* __restrict__ packed, const uint64_t* __restrict__ src_off,
                                 float* const* __restrict__ dst, const uint32_t* __restrict__ len) {
  const int i = blockIdx.x;
  const float* s_ = packed + src_off[i];
  float* d_ = dst[i];
  const uint32_t L = len[i];
  for (uint32_t k = blockIdx.y * blockDim.x + threadIdx.x; k < L; k += gridDim.y * blockDim.x) d_[k] = s_[k];
}

// The same for FP64 rows (observation spline values per slot).
__global__ void k_scatter_rows64(double* __restrict__ dst, size_t stride, const double* __restrict__ packed,
                                 const int* __restrict__ slot) {
  const int i = blockIdx.x;
  const double* s_ = packed + (size_t)i * stride;
  double* d_ = dst + (size_t)slot[i] * stride;
  for (size_t k = blockIdx.y * blockDim.x + threadIdx.x; k < stride; k += (size_t)gridDim.y * blockDim.x) d_[k] = s_[k];
}

// out[b][s - t0] = sum over cells of transnum[s][tr][cell]; one workgroup per (b, s).
__global__ void __launch_bounds__(256) k_sum_transnum(const int* __restrict__ inst, const int* __restrict__ tn,
                                                      size_t traj_sz, int AG, int T, int tr, int t0, int nsteps,
                                                      double* __restrict__ out) {
  __shared__ double red[8];
  const int b = blockIdx.y, s = t0 + blockIdx.x;
  const int* p = tn + (size_t)inst[b] * traj_sz + ((size_t)s * T + tr) * AG;
  double v = 0;
  for (int c = threadIdx.x; c < AG; c += blockDim.x) v += p[c];
  double t = block_sum(v, red);
  if (threadIdx.x == 0) out[(size_t)b * nsteps + blockIdx.x] = t;
}

// ---- diagnostics: raw generator and sampler outputs for the tests ----

__global__ void k_dbg_philox(int n, const u4* __restrict__ ctr, uint32_t k0, uint32_t k1, u4* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = philox4x32_10(ctr[i], k0, k1);
}

// Draw i is the simulation kernels' queued large draw (draw_item: its own
// stream) of cell i, transition 0, at step 0 of instance key, so these
// samplers are tested with the stream numbering the kernels use.
__global__ void k_dbg_poisson(int n, const float* __restrict__ lam, uint64_t seed, uint32_t key, int* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = lam[i] > 0.f ? draw_item(ITEM_POIS, lam[i], 0, (uint32_t)seed, (uint32_t)(seed >> 32), 0u, (uint32_t)i, 0u,
                                    key)
                        : 0;
}

__global__ void k_dbg_binomial(int n, const int* __restrict__ nn, const double* __restrict__ p, uint64_t seed,
                               uint32_t key, int* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = draw_item(ITEM_BINOM, (float)p[i], nn[i], (uint32_t)seed, (uint32_t)(seed >> 32), 0u, (uint32_t)i, 0u, key);
}

}  // namespace beepdev
