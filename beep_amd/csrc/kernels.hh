// Device kernels of the tau-leap simulation / likelihood / MBP path.
//
// Data layout in HBM (per instance, cell = area*G + dp fastest so a wavefront
// touches consecutive addresses):
//   state      [C][AG]     int32  current populations
//   transnum   [S][T][AG]  int32  realised transitions
//   transmean  [S][T][AG]  fp32   Poisson means
//   poptraj    [S][C][AG]  int32  populations at the start of each step
// Per parameter slot (fp32, one block): rate[T][G], sus[G], inf[C], N[G][G]
// (or [S][G][G]), efoi[S], geo[S], beta[S], and af[period][A] when the area
// factor depends on the parameters (else one table per context); the efoi age
// distribution [G] is the context's.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fused_common.hh"

namespace beepdev {

// One workgroup advances one instance over steps [t0, t1).  Each thread owns
// the cells tid, tid + blockDim, ...; populations live in LDS for the whole
// horizon and only the trajectory (transnum, transmean, pop) goes to HBM.
// Per step: (1) infectivity per cell, (2) area mixing, (3a) means, small
// draws in place, large draws queued, (3b) queued draws by all lanes.
template <int MODE>
__global__ void __launch_bounds__(BEEP_FUSED_MAX_BLOCK, 4)
    k_fused(const Job* __restrict__ jobs, uint64_t seed, uint32_t stream) {
  const Job jb = jobs[blockIdx.x];
  // MODE_ANY launches mix simulations and proposals: the job says which.
  const int mode = MODE == MODE_ANY ? jb.mode : MODE;
  ctopo* __restrict__ tpp = (ctopo*)jb.topo;
  const int A = tpp->A, G = tpp->G, C = tpp->C, T = tpp->T, S = tpp->S, AG = tpp->AG, inf_tr = tpp->inf_tr;
  const int nt = blockDim.x, tid = threadIdx.x;
  const float dt = tpp->dt;
  const int qcap = fused_qcap(AG);

  extern __shared__ int smem[];
  int* popA = smem;                                    // [C][AG]
  int* popB = popA + (size_t)C * AG;                   // [C][AG]
  float* sdinf = (float*)(popB + (size_t)C * AG);      // [AG]
  float* sI = sdinf + AG;                              // [AG]
  uint32_t* qcode = (uint32_t*)(sI + AG);              // [qcap]
  float* qval = (float*)(qcode + qcap);                // [qcap]
  int* qcount = (int*)(qval + qcap);                   // [4]
  float* srate = (float*)(qcount + 4);                 // [T][G]
  float* sN = srate + T * G;                           // [G][G]
  float* ssus = sN + G * G;                            // [G]
  float* sinf = ssus + G;                              // [C]
  int* sfrom = (int*)(sinf + C);                       // [T]
  int* sto = sfrom + T;                                // [T]
  gcfloat* diag = gptr(tpp->diag);
  gcfloat* fac = gptr(tpp->fac);
  gcint* in_ptr = gptr(tpp->in_ptr);
  gcint* in_src = gptr(tpp->in_src);
  gcfloat* in_w = gptr(tpp->in_w);
  gcint* order = gptr(tpp->mix_order);
  const bool staged = tpp->stage_mix != 0;
  float* sdiag = (float*)(sto + T);
  float* sfac = sdiag + A;
  int* sptr = (int*)(sfac + A);
  int* ssrc = sptr + A + 1;
  float* sw = (float*)(ssrc + tpp->nnz);
  int* sord = (int*)(sw + tpp->nnz);
  if (staged) {
    for (int i = tid; i < A; i += nt) { sdiag[i] = diag[i]; sfac[i] = fac[i]; sord[i] = order[i]; }
    for (int i = tid; i <= A; i += nt) sptr[i] = in_ptr[i];
    for (int i = tid; i < tpp->nnz; i += nt) { ssrc[i] = in_src[i]; sw[i] = in_w[i]; }
  }

  gcfloat* __restrict__ tab = gptr(jb.tab);
  {
    gcint* src = gptr(jb.from_init ? tpp->pop_init : jb.state);
    for (int i = tid; i < C * AG; i += nt) popA[i] = src[i];
    for (int i = tid; i < T * G; i += nt) srate[i] = tab[tpp->off_rate + i];
    for (int i = tid; i < G * G; i += nt) sN[i] = tab[tpp->off_N + i];
    for (int i = tid; i < G; i += nt) ssus[i] = tab[tpp->off_sus + i];
    for (int i = tid; i < C; i += nt) sinf[i] = tab[tpp->off_inf + i];
    for (int i = tid; i < T; i += nt) { sfrom[i] = gptr(tpp->from)[i]; sto[i] = gptr(tpp->to)[i]; }
    if (tid == 0) qcount[0] = 0;
  }
  __syncthreads();

  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  const size_t TAG = (size_t)T * AG, CAG = (size_t)C * AG;
  int* cur = popA;
  int* nxt = popB;

  for (int s = jb.t0; s < jb.t1; s++) {
    // Phase 1: infectivity of each cell, dinf = sum_c inf_c * pop_c (State::set_I_from_pop).
    for (int cell = tid; cell < AG; cell += nt) {
      float d = 0.f;
      for (int c = 0; c < C; c++) {
        float w = sinf[c];
        if (w != 0.f) d += w * (float)cur[c * AG + cell];
      }
      sdinf[cell] = d;
    }
    __syncthreads();
    // Phase 2: area-mixing contraction Imap = M_off^T dinf, Idiag = diag * dinf,
    // and I = (Idiag + Imap) m + fac Idiag (1 - m)  (State::get_NMI, first half).
    // A time-varying age matrix: this step's.  Step s-1's reads of sN in
    // phase 3a ended at the barrier after it; these writes are seen after
    // the barrier that ends phase 2.
    if (tpp->N_tv) {
      for (int i = tid; i < G * G; i += nt) sN[i] = tab[tpp->off_N + (size_t)s * G * G + i];
    }
    const float m = tab[tpp->off_geo + s];
    if (staged)
      mix_cells((lcfloat*)sdiag, (lcfloat*)sfac, (lcint*)sptr, (lcint*)ssrc, (lcfloat*)sw, (lcint*)sord, sdinf, sI, m,
                G, AG, tid, nt);
    else
      mix_cells(diag, fac, in_ptr, in_src, in_w, order, sdinf, sI, m, G, AG, tid, nt);
    __syncthreads();
    // Phase 3a: age mixing, transition means (State::set_transmean), small
    // draws and the population update (State::update_pop) into the other buffer.
    const float beta_s = tab[tpp->off_beta + s];
    gcfloat* af_s = af_row(tpp, tab, s);
    const float efoi_s = tab[tpp->off_efoi + s];
    gcfloat* agedist = gptr(tpp->agedist);
    int* tn_s = jb.tn ? jb.tn + (size_t)s * TAG : nullptr;
    float* tm_s = jb.tm ? jb.tm + (size_t)s * TAG : nullptr;
    int* pt_s = jb.pt ? jb.pt + (size_t)s * CAG : nullptr;
    gcint* itn = mode == MODE_MBP ? gptr(jb.init_tn) + (size_t)s * TAG : nullptr;
    gcfloat* itm = mode == MODE_MBP ? gptr(jb.init_tm) + (size_t)s * TAG : nullptr;
    gcint* tn_r = mode == MODE_REPLAY ? gptr(jb.tn) + (size_t)s * TAG : nullptr;
    gcbyte* simu_s = mode == MODE_MBP && jb.simu ? gptr(jb.simu) + (size_t)s * A : nullptr;
    const bool update = s < S - 1;           // the last step is not applied (src/state.cc:634)
    const uint32_t c0 = (uint32_t)s | (stream << 16);
    for (int cell = tid; cell < AG; cell += nt) {
      int j = cell / G, g = cell - j * G;
      for (int c = 0; c < C; c++) {
        int v = cur[c * AG + cell];
        nxt[c * AG + cell] = v;
        if (pt_s) st_nt(pt_s, c * AG + cell, v);
      }
      float nmi = 0.f;
      for (int gg = 0; gg < G; gg++) nmi += sN[g * G + gg] * sI[j * G + gg];
      if (nmi < 0.f) nmi = 0.f;
      float afc = af_s[j], agc = agedist[g];
      mod_foi_factors(tpp, s, cell, afc, agc);
      const float foi = foi_mean(beta_s, afc, nmi, efoi_s, agc);
      const bool simu = mode == MODE_MBP ? (simu_s ? simu_s[j] != 0 : false) : true;
      CellRng rng;
      rng.init(k0, k1, c0, (uint32_t)cell, jb.key);
      if (mode != MODE_REPLAY) rng.prime();
#pragma unroll 1
      for (int tr = 0; tr < T; tr++) {
        const int fr = sfrom[tr];
        float popu = (float)cur[fr * AG + cell];
        float lp = 0.f;
        if (popu > 0.f) lp = tr == inf_tr ? dt * popu * ssus[g] * foi : dt * popu * srate[tr * G + g];
        lp = mod_mean(tpp, s, tr, cell, lp);
        // src/state.cc:429 (NaN mean) and src/utils.cc:150 (mean > 1e9) stop the reference.
        if (lp != lp) { atomicOr(jb.err, 1); lp = 0.f; }
        else if (lp > 1e9f) { atomicOr(jb.err, 2); lp = 0.f; }
        if (tm_s) st_nt(tm_s, tr * AG + cell, lp);
        int n = 0;
        int kind = -1;
        float val = 0.f;
        if (mode == MODE_REPLAY) {
          n = tn_r[tr * AG + cell];
        } else if (lp > 0.f) {
          if (mode == MODE_SIM || simu) {
            if (lp < BEEP_POIS_INLINE) n = poisson_inversion(lp, rng.next());
            else { kind = ITEM_POIS; val = lp; }
          } else {
            // Coupled proposal (src/mbp.cc:286-311).
            int ni = itn[tr * AG + cell];
            float li = itm[tr * AG + cell];
            if (lp > li) {
              float dl = lp - li;
              if (dl < BEEP_POIS_INLINE) n = ni + poisson_inversion(dl, rng.next());
              else { kind = ITEM_POIS_ADD; val = dl; }
            } else if (lp == li) {
              n = ni;
            } else {
              float p = lp / li;
              float pp = p > 0.5f ? 1.f - p : p;
              if ((float)ni * pp < 10.f) n = binomial_draw(ni, p, rng.next(), rng);
              else { kind = ITEM_BINOM; val = p; }
            }
          }
        }
        if (kind >= 0) {
          int slot = atomicAdd(qcount, 1);
          if (slot < qcap) {
            qcode[slot] = (uint32_t)cell | ((uint32_t)tr << 20) | ((uint32_t)kind << 28);
            qval[slot] = val;
            continue;                        // drawn in phase 3b
          }
          int ni = mode == MODE_MBP ? itn[tr * AG + cell] : 0;
          n = draw_item(kind, val, ni, k0, k1, c0, (uint32_t)cell, (uint32_t)tr, jb.key);
        }
        if (tn_s && mode != MODE_REPLAY) st_nt(tn_s, tr * AG + cell, n);
        if (update && n != 0) {
          nxt[fr * AG + cell] -= n;
          nxt[sto[tr] * AG + cell] += n;
        }
      }
    }
    __syncthreads();
    // Phase 3b: the queued large draws (PTRS / BTRD), every lane busy.
    if (mode != MODE_REPLAY) {
      int nq = qcount[0];
      if (nq > qcap) nq = qcap;
      for (int i = tid; i < nq; i += nt) {
        uint32_t code = qcode[i];
        int cell = (int)(code & 0xFFFFFu), tr = (int)((code >> 20) & 0xFFu), kind = (int)(code >> 28);
        int ni = mode == MODE_MBP ? itn[tr * AG + cell] : 0;
        int n = draw_item(kind, qval[i], ni, k0, k1, c0, (uint32_t)cell, (uint32_t)tr, jb.key);
        if (tn_s) st_nt(tn_s, tr * AG + cell, n);
        if (update && n != 0) {
          atomicSub(&nxt[sfrom[tr] * AG + cell], n);
          atomicAdd(&nxt[sto[tr] * AG + cell], n);
        }
      }
      __syncthreads();
      if (tid == 0) qcount[0] = 0;
    }
    int* t = cur;
    cur = nxt;
    nxt = t;
  }
  __syncthreads();
  for (int cell = tid; cell < AG; cell += nt)
    for (int c = 0; c < C; c++) ((gint*)jb.state)[c * AG + cell] = cur[c * AG + cell];
}

}  // namespace beepdev
