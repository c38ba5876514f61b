// Structure-specialised transition kernel of the step-wise path, compiled at
// run time together with fused_spec.hh (same prelude: namespace beepspec).
// It computes exactly what k_step_trans (step_kernels.hh) computes -- same
// FP32 expressions and order, same uniforms in the same order -- with the
// cell's populations in registers at compile-time indices, the transition
// loop unrolled, the slot's small tables staged in LDS, and a queue with room
// for every (cell, transition) of the workgroup.
#pragma once

namespace beepdev {

template <int MODE>
__device__ inline void step_spec_body(const Job* __restrict__ jobs, const float* __restrict__ Ibuf,
                                      float* __restrict__ D, int Apad, int s, const uint64_t* __restrict__ rngp) {
  using namespace beepspec;
  const uint64_t seed = rngp[0];
  const uint32_t stream = (uint32_t)rngp[1];
  const int b = blockIdx.x;
  const Job jb = jobs[b];
  ctopo* __restrict__ tp = (ctopo*)jb.topo;
  const int A = tp->A, S = tp->S, AG = tp->AG;
  const float dt = tp->dt;
  const int tid = threadIdx.x;
  const int cell = blockIdx.y * BEEP_STEP_BLOCK + tid;
  const bool valid = cell < AG && s >= jb.t0 && s < jb.t1;

  __shared__ int spop[CP][BEEP_STEP_BLOCK];          // populations after the step's own-lane draws
  __shared__ uint32_t qcode[T * BEEP_STEP_BLOCK];
  __shared__ float qval[T * BEEP_STEP_BLOCK];
  __shared__ int qcount;
  __shared__ float srate[T * G], sN[G * G], ssus[G], sinf[C];
  gcfloat* tab = gptr(jb.tab);
  for (int i = tid; i < T * G; i += BEEP_STEP_BLOCK) srate[i] = tab[tp->off_rate + i];
  const size_t offN = tp->off_N + (tp->N_tv ? (size_t)s * G * G : 0);   // a time-varying age matrix: step s's
  for (int i = tid; i < G * G; i += BEEP_STEP_BLOCK) sN[i] = tab[offN + i];
  for (int i = tid; i < G; i += BEEP_STEP_BLOCK) ssus[i] = tab[tp->off_sus + i];
  for (int i = tid; i < C; i += BEEP_STEP_BLOCK) sinf[i] = tab[tp->off_inf + i];
  if (tid == 0) qcount = 0;
  __syncthreads();

  const size_t TAG = (size_t)T * AG, CAG = (size_t)C * AG;
  int* tn_s = jb.tn ? jb.tn + (size_t)s * TAG : nullptr;
  float* tm_s = jb.tm ? jb.tm + (size_t)s * TAG : nullptr;
  int* pt_s = jb.pt ? jb.pt + (size_t)s * CAG : nullptr;
  gcint* itn = MODE == MODE_MBP ? gptr(jb.init_tn) + (size_t)s * TAG : nullptr;
  gcfloat* itm = MODE == MODE_MBP ? gptr(jb.init_tm) + (size_t)s * TAG : nullptr;
  gcint* tn_r = MODE == MODE_REPLAY ? gptr(jb.tn) + (size_t)s * TAG : nullptr;
  gint* state = (gint*)jb.state;
  const bool update = s < S - 1;
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  const uint32_t c0 = (uint32_t)s | (stream << 16);
  int j = 0, g = 0;
  if (valid) {
    j = cell / G;
    g = cell - j * G;
    int p[CP];
#pragma unroll
    for (int c = 0; c < CP; c++) p[c] = c < C ? state[c * AG + cell] : 0;
    if (pt_s) {
#pragma unroll
      for (int c = 0; c < C; c++) st_nt(pt_s, c * AG + cell, p[c]);
    }
    int q[CP];
#pragma unroll
    for (int c = 0; c < CP; c++) q[c] = p[c];
    gcfloat* Ib = gptr(Ibuf) + (size_t)b * G * Apad;
    float nmi = 0.f;
#pragma unroll
    for (int gg = 0; gg < G; gg++) nmi += sN[g * G + gg] * Ib[(size_t)gg * Apad + j];
    if (nmi < 0.f) nmi = 0.f;
    float bet, ef, agc;
    foi_groups(tp, tab, s, j, g, DPM, bet, ef, agc);
    float afc = af_row(tp, tab, s)[j];
    mod_foi_factors(tp, s, cell, afc, agc);
    const float foi = foi_mean(bet, afc, nmi, ef, agc);
    const bool simu = MODE == MODE_MBP ? (jb.simu ? gptr(jb.simu)[(size_t)s * A + j] != 0 : false) : true;
    CellRng rng;
    rng.init(k0, k1, c0, (uint32_t)cell, jb.key);
    if (MODE != MODE_REPLAY) rng.prime();
#pragma unroll
    for (int tr = 0; tr < T; tr++) {
      const int fr = FROM[tr], to = TO[tr];
      float popu = (float)p[fr];
      float lp = 0.f;
      if (popu > 0.f) lp = tr == INF_TR ? dt * popu * ssus[g] * foi : dt * popu * srate[tr * G + g];
      lp = mod_mean(tp, s, tr, cell, lp);
      if (lp != lp) { atomicOr(jb.err, 1); lp = 0.f; }
      else if (lp > 1e9f) { atomicOr(jb.err, 2); lp = 0.f; }
      if (tm_s) st_nt(tm_s, tr * AG + cell, lp);
      int n = 0, kind = -1;
      float val = 0.f;
      if (MODE == MODE_REPLAY) {
        n = tn_r[tr * AG + cell];
      } else if (lp > 0.f) {
        if (MODE == MODE_SIM || simu) {
          if (lp < BEEP_POIS_INLINE) n = poisson_inversion(lp, rng.next());
          else { kind = ITEM_POIS; val = lp; }
        } else {
          int ni = itn[tr * AG + cell];
          float li = itm[tr * AG + cell];
          if (lp > li) {
            float dl = lp - li;
            if (dl < BEEP_POIS_INLINE) n = ni + poisson_inversion(dl, rng.next());
            else { kind = ITEM_POIS_ADD; val = dl; }
          } else if (lp == li) {
            n = ni;
          } else {
            float pr = lp / li;
            float pp = pr > 0.5f ? 1.f - pr : pr;
            if ((float)ni * pp < 10.f) n = binomial_draw_small(ni, pr, rng.next(), rng);
            else { kind = ITEM_BINOM; val = pr; }
          }
        }
      }
      if (kind >= 0) {
        int slot = atomicAdd(&qcount, 1);
        qcode[slot] = (uint32_t)tid | ((uint32_t)tr << 20) | ((uint32_t)kind << 28);
        qval[slot] = val;
      } else {
        if (tn_s && MODE != MODE_REPLAY) st_nt(tn_s, tr * AG + cell, n);
        if (update) {
          q[fr] -= n;
          q[to] += n;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < CP; c++) spop[c][tid] = q[c];
  }
  __syncthreads();
  const int nq = qcount;
  for (int i = tid; i < nq; i += BEEP_STEP_BLOCK) {
    uint32_t code = qcode[i];
    int t2 = (int)(code & 0xFFFFFu), tr = (int)((code >> 20) & 0xFFu), kind = (int)(code >> 28);
    int c2 = blockIdx.y * BEEP_STEP_BLOCK + t2;
    int ni = MODE == MODE_MBP ? itn[tr * AG + c2] : 0;
    int n = draw_item_inline(kind, qval[i], ni, k0, k1, c0, (uint32_t)c2, (uint32_t)tr, jb.key);
    if (tn_s) st_nt(tn_s, tr * AG + c2, n);
    if (update && n != 0) {
      int fr = 0, to = 0;
#pragma unroll
      for (int k = 0; k < T; k++)
        if (k == tr) { fr = FROM[k]; to = TO[k]; }
      atomicSub(&spop[fr][t2], n);
      atomicAdd(&spop[to][t2], n);
    }
  }
  __syncthreads();
  if (valid) {
    float d = 0.f;
#pragma unroll
    for (int c = 0; c < C; c++) {
      int v = spop[c][tid];
      state[c * AG + cell] = v;
      if (INFC[c]) {
        float w = sinf[c];
        if (w != 0.f) d += w * (float)v;
      }
    }
    ((gfloat*)D)[frag_at(b * G + g, j, Apad)] = d;
  }
}

}  // namespace beepdev

extern "C" __global__ void __launch_bounds__(BEEP_STEP_BLOCK, beepspec::STEP_WAVES)
    beep_step_spec_sim(const beepdev::Job* jobs, const float* I, float* D, int Apad, int s, const uint64_t* rngp) {
  beepdev::step_spec_body<beepdev::MODE_SIM>(jobs, I, D, Apad, s, rngp);
}
extern "C" __global__ void __launch_bounds__(BEEP_STEP_BLOCK, beepspec::STEP_WAVES)
    beep_step_spec_mbp(const beepdev::Job* jobs, const float* I, float* D, int Apad, int s, const uint64_t* rngp) {
  beepdev::step_spec_body<beepdev::MODE_MBP>(jobs, I, D, Apad, s, rngp);
}
extern "C" __global__ void __launch_bounds__(BEEP_STEP_BLOCK, beepspec::STEP_WAVES)
    beep_step_spec_replay(const beepdev::Job* jobs, const float* I, float* D, int Apad, int s, const uint64_t* rngp) {
  beepdev::step_spec_body<beepdev::MODE_REPLAY>(jobs, I, D, Apad, s, rngp);
}
