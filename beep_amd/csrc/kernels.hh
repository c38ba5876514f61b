// Device kernels of the tau-leap simulation / likelihood / MBP path.
//
// Data layout in HBM (per instance, cell = area*G + dp fastest so a wavefront
// touches consecutive addresses):
//   state      [C][AG]     int32  current populations
//   transnum   [S][T][AG]  int32  realised transitions
//   transmean  [S][T][AG]  fp32   Poisson means
//   poptraj    [S][C][AG]  int32  populations at the start of each step
// Per parameter slot (fp32, one block): rate[T][G], sus[G], inf[C], N[G][G],
// efoi[S][G], geo[S], beta_af[S][A].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_rng.hh"

namespace beepdev {

struct TopoDev {
  int A, G, C, T, S, AG, inf_tr, spd;
  float dt;
  int off_rate, off_sus, off_inf, off_N, off_efoi, off_geo, off_beta;
  const int* from;       // [T]
  const int* to;         // [T]
  const float* diag;     // [A]  normalised diagonal of the area mixing
  const float* fac;      // [A]  genQ.factor
  const int* in_ptr;     // [A+1] incoming off-diagonal mixing (gather form)
  const int* in_src;     // [nnz]
  const float* in_w;     // [nnz]
  const int* pop_init;   // [C][AG]
};

// One instance's work in a launch.
struct Job {
  const TopoDev* topo;
  const float* tab;           // parameter slot tables
  int* state;                 // [C][AG]
  int* tn;                    // [S][T][AG] or null
  float* tm;                  // [S][T][AG] or null
  int* pt;                    // [S][C][AG] or null
  const int* init_tn;         // MBP: initial instance's transnum
  const float* init_tm;       // MBP: initial instance's transmean
  const uint8_t* simu;        // MBP: [S][A] 1 = simulate afresh, or null
  int* err;                   // bit 0: NaN mean, bit 1: mean > 1e9
  uint32_t key;
  int t0, t1;
  int from_init;              // 1: start from topo->pop_init
};

enum { MODE_SIM = 0, MODE_MBP = 1, MODE_REPLAY = 2 };

// LDS bytes a job needs.
__host__ __device__ inline size_t fused_lds_bytes(int AG, int C, int T, int G) {
  return (size_t)4 * ((size_t)C * AG + 2 * (size_t)AG + (size_t)T * G + (size_t)G * G + G + C + 2 * T);
}

template <int MODE, int TMAX>
__global__ void __launch_bounds__(256) k_fused(const Job* __restrict__ jobs, uint64_t seed, uint32_t stream) {
  const Job jb = jobs[blockIdx.x];
  const TopoDev tp = *jb.topo;
  const int A = tp.A, G = tp.G, C = tp.C, T = tp.T, S = tp.S, AG = tp.AG, inf_tr = tp.inf_tr;
  const int nt = blockDim.x, tid = threadIdx.x;
  const float dt = tp.dt;

  extern __shared__ int smem[];
  int* spop = smem;                                    // [C][AG]
  float* sdinf = (float*)(spop + (size_t)C * AG);      // [AG]
  float* sI = sdinf + AG;                              // [AG]
  float* srate = sI + AG;                              // [T][G]
  float* sN = srate + T * G;                           // [G][G]
  float* ssus = sN + G * G;                            // [G]
  float* sinf = ssus + G;                              // [C]
  int* sfrom = (int*)(sinf + C);                       // [T]
  int* sto = sfrom + T;                                // [T]

  const float* tab = jb.tab;
  const int* src = jb.from_init ? tp.pop_init : jb.state;
  for (int i = tid; i < C * AG; i += nt) spop[i] = src[i];
  for (int i = tid; i < T * G; i += nt) srate[i] = tab[tp.off_rate + i];
  for (int i = tid; i < G * G; i += nt) sN[i] = tab[tp.off_N + i];
  for (int i = tid; i < G; i += nt) ssus[i] = tab[tp.off_sus + i];
  for (int i = tid; i < C; i += nt) sinf[i] = tab[tp.off_inf + i];
  for (int i = tid; i < T; i += nt) { sfrom[i] = tp.from[i]; sto[i] = tp.to[i]; }
  __syncthreads();

  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  const size_t TAG = (size_t)T * AG, CAG = (size_t)C * AG;

  for (int s = jb.t0; s < jb.t1; s++) {
    // Phase 1: infectivity of each cell, dinf = sum_c inf_c * pop_c (State::set_I_from_pop).
    for (int cell = tid; cell < AG; cell += nt) {
      float d = 0.f;
      for (int c = 0; c < C; c++) {
        float w = sinf[c];
        if (w != 0.f) d += w * (float)spop[c * AG + cell];
      }
      sdinf[cell] = d;
    }
    __syncthreads();
    // Phase 2: area-mixing contraction Imap = M_off^T dinf, Idiag = diag * dinf,
    // and I = (Idiag + Imap) m + fac Idiag (1 - m)  (State::get_NMI first half).
    const float m = tab[tp.off_geo + s];
    for (int cell = tid; cell < AG; cell += nt) {
      int j = cell / G, g = cell - j * G;
      float idg = tp.diag[j] * sdinf[cell];
      float imp = 0.f;
      int e1 = tp.in_ptr[j + 1];
      for (int e = tp.in_ptr[j]; e < e1; e++) imp += tp.in_w[e] * sdinf[tp.in_src[e] * G + g];
      sI[cell] = (idg + imp) * m + tp.fac[j] * idg * (1.f - m);
    }
    __syncthreads();
    // Phase 3: age mixing, transition means, draws and the population update.
    const float* beta_s = tab + tp.off_beta + (size_t)s * A;
    const float* efoi_s = tab + tp.off_efoi + (size_t)s * G;
    int* tn_s = jb.tn ? jb.tn + (size_t)s * TAG : nullptr;
    float* tm_s = jb.tm ? jb.tm + (size_t)s * TAG : nullptr;
    for (int cell = tid; cell < AG; cell += nt) {
      int j = cell / G, g = cell - j * G;
      float nmi = 0.f;
      for (int gg = 0; gg < G; gg++) nmi += sN[g * G + gg] * sI[j * G + gg];
      if (nmi < 0.f) nmi = 0.f;
      const float be = beta_s[j];
      const float et = efoi_s[g];
      float lam[TMAX];
      int nn[TMAX];
#pragma unroll
      for (int tr = 0; tr < TMAX; tr++) {
        lam[tr] = 0.f;
        nn[tr] = 0;
        if (tr < T) {
          float popu = (float)spop[sfrom[tr] * AG + cell];
          if (popu > 0.f)
            lam[tr] = tr == inf_tr ? dt * popu * ssus[g] * (be * nmi + et) : dt * popu * srate[tr * G + g];
          // src/state.cc:429 (NaN mean) and src/utils.cc:150 (mean > 1e9) abort the reference.
          if (lam[tr] != lam[tr]) { atomicOr(jb.err, 1); lam[tr] = 0.f; }
          else if (lam[tr] > 1e9f) { atomicOr(jb.err, 2); lam[tr] = 0.f; }
        }
      }
      if (MODE == MODE_REPLAY) {
#pragma unroll
        for (int tr = 0; tr < TMAX; tr++)
          if (tr < T) nn[tr] = tn_s[tr * AG + cell];
      } else {
        const int* itn = MODE == MODE_MBP ? jb.init_tn + (size_t)s * TAG : nullptr;
        const float* itm = MODE == MODE_MBP ? jb.init_tm + (size_t)s * TAG : nullptr;
        const bool simu = MODE == MODE_MBP ? (jb.simu ? jb.simu[(size_t)s * A + j] != 0 : false) : true;
        const uint32_t c0 = (uint32_t)s | (stream << 16);
#pragma unroll
        for (int q = 0; q < (TMAX + 3) / 4; q++) {
          bool need = false;
#pragma unroll
          for (int r = 0; r < 4; r++) {
            int tr = 4 * q + r;
            if (tr < TMAX && tr < T && lam[tr] > 0.f) need = true;
          }
          if (!need) continue;
          u4 blk = philox4x32_10(u4{c0, (uint32_t)cell, (uint32_t)q, jb.key}, k0, k1);
#pragma unroll
          for (int r = 0; r < 4; r++) {
            int tr = 4 * q + r;
            if (tr >= TMAX || tr >= T) continue;
            float lp = lam[tr];
            uint32_t bits = pick(blk, r);
            ExtraUniforms xu{k0, k1, c0, (uint32_t)cell, jb.key, (uint32_t)tr, 0u, 4, u4{0, 0, 0, 0}};
            int n = 0;
            if (MODE == MODE_SIM) {
              if (lp > 0.f) n = poisson_draw(lp, u01f(bits), u01d(bits), xu);
            } else {
              int ni = itn[tr * AG + cell];
              if (lp > 0.f) {
                if (simu) n = poisson_draw(lp, u01f(bits), u01d(bits), xu);
                else {
                  float li = itm[tr * AG + cell];
                  if (lp > li) {
                    float dl = lp - li;
                    n = ni + poisson_draw(dl, u01f(bits), u01d(bits), xu);
                  } else if (lp == li) n = ni;
                  else n = binomial_draw(ni, (double)lp / (double)li, u01d(bits), xu);
                }
              }
            }
            nn[tr] = n;
          }
        }
        if (MODE == MODE_MBP) {
          // Cells whose proposed mean is zero draw nothing (src/mbp.cc:289).
#pragma unroll
          for (int tr = 0; tr < TMAX; tr++)
            if (tr < T && !(lam[tr] > 0.f)) nn[tr] = 0;
        }
      }
      // Outputs of this step.
      if (jb.pt) {
        int* pt_s = jb.pt + (size_t)s * CAG;
        for (int c = 0; c < C; c++) pt_s[c * AG + cell] = spop[c * AG + cell];
      }
#pragma unroll
      for (int tr = 0; tr < TMAX; tr++) {
        if (tr < T) {
          if (tn_s && MODE != MODE_REPLAY) tn_s[tr * AG + cell] = nn[tr];
          if (tm_s) tm_s[tr * AG + cell] = lam[tr];
        }
      }
      // State::update_pop: the last step of the horizon is not applied.
      if (s < S - 1) {
#pragma unroll
        for (int tr = 0; tr < TMAX; tr++) {
          if (tr < T && nn[tr] != 0) {
            spop[sfrom[tr] * AG + cell] -= nn[tr];
            spop[sto[tr] * AG + cell] += nn[tr];
          }
        }
      }
    }
    // Phase 1 of the next step writes sdinf, which other threads read in phase
    // 2 of this step; the barrier after phase 2 already orders those accesses,
    // and each thread only touches the populations of its own cells.
  }
  for (int cell = tid; cell < AG; cell += nt)
    for (int c = 0; c < C; c++) jb.state[c * AG + cell] = spop[c * AG + cell];
}

}  // namespace beepdev
