// Structure-specialised fused simulation kernel, compiled at run time for one
// compartment structure (hiprtc).  The text compiled is a prelude defining
// namespace beepspec, then device_rng.hh, fused_common.hh and this file.  The
// prelude fixes:
//   constexpr int C, T, G, CP;        compartments, transitions, age groups,
//                                     C rounded up to a multiple of 4
//   constexpr int INF_TR;             the infection transition
//   constexpr int FROM[T], TO[T];     transition source / destination
//   constexpr bool INFC[C];           compartments with an infectivity weight
//   constexpr int NSTRAIN, DPM;       strains, and groups per strain (G = NSTRAIN * DPM)
//
// It computes exactly what k_fused (kernels.hh) computes -- the same FP32
// expressions in the same order and the same uniforms consumed in the same
// order -- so trajectories are bit-identical to the generic kernel and to the
// step-wise path for the same keys.  What the specialisation buys: a cell's
// populations sit in registers through the step with compile-time source and
// destination indices, the transition loop is unrolled, divisions by G are by
// a constant, and populations are single-buffered in LDS in 16-byte chunks of
// four compartments so one cell's compartments load and store as vectors.
//
// Per step: (B) area mixing, (C) means, small draws, updates of the cell's own
// populations, large draws queued, (D) the queued draws by all lanes, applied
// with LDS atomics, (E) the infectivity of each cell for the next step.
#pragma once

namespace beepdev {

static_assert(beepspec::CP % 4 == 0 && beepspec::CP >= beepspec::C, "CP");

// Populations in LDS: compartment c of a cell in 16-byte chunk c / 4 of the
// cell; chunk k of all cells is contiguous, so one wavefront's 16-byte
// accesses to its cells' chunk are consecutive in LDS.
__device__ inline int pidx(int c, int cell, int AG) { return ((c >> 2) * AG + cell) * 4 + (c & 3); }

// Strains: for each strain-0 cell, the susceptibles of its dp summed over
// the strains (reads pop only; the caller's barrier publishes spool).
__device__ inline void pool_susceptibles(const int* pop, int* spool, int AG, int tid, int nt) {
  if constexpr (beepspec::NSTRAIN > 1) {
    constexpr int G = beepspec::G, DPM = beepspec::DPM, SUS = beepspec::FROM[beepspec::INF_TR];
    for (int cell = tid; cell < AG; cell += nt) {
      const int j = cell / G, g = cell - j * G;
      if (g < DPM) {
        int sum = 0;
#pragma unroll
        for (int st = 0; st < beepspec::NSTRAIN; st++) sum += pop[pidx(SUS, j * G + st * DPM + g, AG)];
        spool[cell] = sum;
      }
    }
  }
}

template <int MODE>
__device__ inline void fused_spec_body(const Job* __restrict__ jobs, uint64_t seed, uint32_t stream) {
  using namespace beepspec;
  const Job jb = jobs[blockIdx.x];
  const int mode = MODE == MODE_ANY ? jb.mode : MODE;
  ctopo* __restrict__ tpp = (ctopo*)jb.topo;
  const int A = tpp->A, S = tpp->S, AG = tpp->AG;
  const int nt = blockDim.x, tid = threadIdx.x;
  const float dt = tpp->dt;
  const int qcap = T * AG;                             // every (cell, transition): never full
  const bool q16 = spec_queue16(T, AG);

  extern __shared__ int smem[];
  int* pop = smem;                                     // [CP/4][AG] chunks of 4 compartments
  float* sdinf = (float*)(pop + (size_t)AG * CP);      // [AG]
  float* sI = sdinf + AG;                              // [AG]
  float* qval = sI + AG;                               // [qcap] value; negative: ITEM_POIS_ADD
  int* qcount = (int*)(qval + qcap);                   // [4]
  float* srate = (float*)(qcount + 4);                 // [T][G]
  float* sN = srate + T * G;                           // [G][G]
  float* ssus = sN + G * G;                            // [G]
  float* sinf = ssus + G;                              // [C]
  gcfloat* diag = gptr(tpp->diag);
  gcfloat* fac = gptr(tpp->fac);
  gcint* in_ptr = gptr(tpp->in_ptr);
  gcint* in_src = gptr(tpp->in_src);
  gcfloat* in_w = gptr(tpp->in_w);
  gcint* order = gptr(tpp->mix_order);
  const bool staged = tpp->stage_mix != 0;
  // This step's area factor and the external force's age weights, read by
  // every cell in (C): from LDS, so that no global load there waits behind the
  // trajectory stores still in flight (loads and stores share one counter).
  float* saf = sinf + C;                               // [A]
  const int nE = tpp->nE, grouped = tpp->nR > 1 || nE > 1;
  float* sage = saf + A;                               // [nE][G] age weights of each external-force group
  // With R or external-force groups: this step's beta of every area and
  // external force of every (strain, area), and each (strain, area)'s group.
  float* sbeta = sage + (size_t)nE * G;                // [A]
  float* sefoi = sbeta + (grouped ? A : 0);            // [NSTRAIN][A]
  int* segrp = (int*)(sefoi + (grouped ? NSTRAIN * A : 0));   // [NSTRAIN][A]
  float* sdiag = (float*)(segrp + (grouped ? NSTRAIN * A : 0));
  float* sfac = sdiag + A;
  int* sptr = (int*)(sfac + A);
  int* ssrc = sptr + A + 1;
  float* sw = (float*)(ssrc + tpp->nnz);
  int* sord = (int*)(sw + tpp->nnz);
  // The queue's item indices (cell * T + tr), after the tables and, when
  // staged, the mixing CSR: 16 bits each when they fit.
  void* qidx = staged ? (void*)(sord + A) : (void*)sdiag;
  uint16_t* qidx16 = (uint16_t*)qidx;
  uint32_t* qidx32 = (uint32_t*)qidx;
  // Strains: each strain-0 cell's susceptibles pooled over the strains of its
  // dp (src/state.cc:395-404), after the queue's indices.
  int* spool = (int*)((char*)qidx + ((size_t)qcap * (q16 ? 2 : 4) + 3) / 4 * 4);
  if (staged) {
    for (int i = tid; i < A; i += nt) { sdiag[i] = diag[i]; sfac[i] = fac[i]; sord[i] = order[i]; }
    for (int i = tid; i <= A; i += nt) sptr[i] = in_ptr[i];
    for (int i = tid; i < tpp->nnz; i += nt) { ssrc[i] = in_src[i]; sw[i] = in_w[i]; }
  }

  gcfloat* __restrict__ tab = gptr(jb.tab);
  {
    gcint* src = gptr(jb.from_init ? tpp->pop_init : jb.state);
    for (int cell = tid; cell < AG; cell += nt) {
#pragma unroll
      for (int c = 0; c < CP; c++) pop[pidx(c, cell, AG)] = c < C ? src[c * AG + cell] : 0;
    }
    for (int i = tid; i < T * G; i += nt) srate[i] = tab[tpp->off_rate + i];
    for (int i = tid; i < G * G; i += nt) sN[i] = tab[tpp->off_N + i];
    for (int i = tid; i < G; i += nt) ssus[i] = tab[tpp->off_sus + i];
    for (int i = tid; i < C; i += nt) sinf[i] = tab[tpp->off_inf + i];
    for (int i = tid; i < nE * G; i += nt) sage[i] = gptr(tpp->agedist)[i];
    if (grouped)
      for (int i = tid; i < NSTRAIN * A; i += nt) segrp[i] = tpp->nE > 1 ? gptr(tpp->egrp)[i] : 0;
    if (tid == 0) qcount[0] = 0;
  }
  __syncthreads();

  // (E) for the first step: dinf = sum_c inf_c * pop_c (State::set_I_from_pop).
  for (int cell = tid; cell < AG; cell += nt) {
    const int4* pv = (const int4*)pop;
    float d = 0.f;
#pragma unroll
    for (int c4 = 0; c4 < CP / 4; c4++) {
      int4 v = pv[c4 * AG + cell];
      int pc[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int c = 4 * c4 + k;
        if (c < C && INFC[c]) {
          float w = sinf[c];
          if (w != 0.f) d += w * (float)pc[k];
        }
      }
    }
    sdinf[cell] = d;
  }
  pool_susceptibles(pop, spool, AG, tid, nt);
  __syncthreads();

  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  const size_t TAG = (size_t)T * AG, CAG = (size_t)C * AG;

  for (int s = jb.t0; s < jb.t1; s++) {
    const float beta_s = tab[tpp->off_beta + s];
    gcfloat* af_s = af_row(tpp, tab, s);
    const float efoi_s = tab[tpp->off_efoi + s];
    // Read in (C) after the barrier that ends (B); step s-1's reads ended at
    // the barriers of (C), (D) and (E).
    for (int i = tid; i < A; i += nt) saf[i] = af_s[i];
    if (grouped) {
      const int nR = tpp->nR;
      for (int i = tid; i < A; i += nt) sbeta[i] = tab[tpp->off_beta + (size_t)s * nR + (nR > 1 ? gptr(tpp->rgrp)[i] : 0)];
      for (int i = tid; i < NSTRAIN * A; i += nt) sefoi[i] = tab[tpp->off_efoi + (size_t)s * nE + segrp[i]];
    }
    // A time-varying age matrix: this step's.  Step s-1's reads of sN in (C)
    // ended at the barrier after (C); these writes are seen after (B)'s.
    if (tpp->N_tv) {
      for (int i = tid; i < G * G; i += nt) sN[i] = tab[tpp->off_N + (size_t)s * G * G + i];
    }
    // (B) area mixing, from LDS when the CSR is staged there.
    const float m = tab[tpp->off_geo + s];
    if (staged)
      mix_cells((lcfloat*)sdiag, (lcfloat*)sfac, (lcint*)sptr, (lcint*)ssrc, (lcfloat*)sw, (lcint*)sord, sdinf, sI, m,
                G, AG, tid, nt);
    else
      mix_cells(diag, fac, in_ptr, in_src, in_w, order, sdinf, sI, m, G, AG, tid, nt);
    __syncthreads();
    // (C) means, small draws, own-cell updates; large draws queued.
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
      const int j = cell / G, g = cell - j * G;
      // The cell's global reads are issued before its first store: a load
      // issued after a store waits for that store too.
      const bool simu = mode == MODE_MBP ? (simu_s ? simu_s[j] != 0 : false) : true;
      int nrep[T], nini[T];
      float lini[T];
      if (mode == MODE_REPLAY) {
#pragma unroll
        for (int tr = 0; tr < T; tr++) nrep[tr] = tn_r[tr * AG + cell];
      }
      if (mode == MODE_MBP && !simu) {
#pragma unroll
        for (int tr = 0; tr < T; tr++) {
          nini[tr] = itn[tr * AG + cell];
          lini[tr] = itm[tr * AG + cell];
        }
      }
      int p[CP];
      {
        const int4* pv = (const int4*)pop;
#pragma unroll
        for (int c4 = 0; c4 < CP / 4; c4++) {
          int4 v = pv[c4 * AG + cell];
          p[4 * c4] = v.x; p[4 * c4 + 1] = v.y; p[4 * c4 + 2] = v.z; p[4 * c4 + 3] = v.w;
        }
      }
      int pool = 0;
      if constexpr (NSTRAIN > 1) {
        // Strain st's infections draw on the pooled susceptibles, which the
        // reference keeps in strain 0's dp (src/state.cc:395-425).
        const int st = g / DPM;
        pool = spool[j * G + (g - st * DPM)];
        p[FROM[INF_TR]] = st == 0 ? pool : 0;
      }
      if (pt_s) {
#pragma unroll
        for (int c = 0; c < C; c++) st_nt(pt_s, c * AG + cell, p[c]);
      }
      int q[CP];
#pragma unroll
      for (int c = 0; c < CP; c++) q[c] = p[c];
      float nmi = 0.f;
#pragma unroll
      for (int gg = 0; gg < G; gg++) nmi += sN[g * G + gg] * sI[j * G + gg];
      if (nmi < 0.f) nmi = 0.f;
      float afc = saf[j], agc, bet = beta_s, ef = efoi_s;
      if (grouped) {
        const int k = (g / DPM) * A + j;
        bet = sbeta[j];
        ef = sefoi[k];
        agc = sage[segrp[k] * G + g];
      } else {
        agc = sage[g];
      }
      mod_foi_factors(tpp, s, cell, afc, agc);
      const float foi = foi_mean(bet, afc, nmi, ef, agc);
      CellRng rng;
      rng.init(k0, k1, c0, (uint32_t)cell, jb.key);
      if (mode != MODE_REPLAY) rng.prime();
#pragma unroll
      for (int tr = 0; tr < T; tr++) {
        const int fr = FROM[tr], to = TO[tr];
        float popu = (float)p[fr];
        if constexpr (NSTRAIN > 1) {
          if (tr == INF_TR) popu = (float)pool;
        }
        float lp = 0.f;
        if (popu > 0.f) lp = tr == INF_TR ? dt * popu * ssus[g] * foi : dt * popu * srate[tr * G + g];
        lp = mod_mean(tpp, s, tr, cell, lp);
        // src/state.cc:429 (NaN mean) and src/utils.cc:150 (mean > 1e9) stop the reference.
        if (lp != lp) { atomicOr(jb.err, 1); lp = 0.f; }
        else if (lp > 1e9f) { atomicOr(jb.err, 2); lp = 0.f; }
        if (tm_s) st_nt(tm_s, tr * AG + cell, lp);
        int n = 0;
        int kind = -1;
        float val = 0.f;
        if (mode == MODE_REPLAY) {
          n = nrep[tr];
        } else if (lp > 0.f) {
          if (mode == MODE_SIM || simu) {
            if (lp < BEEP_POIS_INLINE) n = poisson_inversion(lp, rng.next());
            else { kind = ITEM_POIS; val = lp; }
          } else {
            // Coupled proposal (src/mbp.cc:286-311).
            const int ni = nini[tr];
            const float li = lini[tr];
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
        // Draws queued for (D): the wavefront's lanes that queue this
        // transition take consecutive slots from one LDS atomic (the draws are
        // keyed by cell and transition, so the slots' order is immaterial).
        const unsigned long long qm = __ballot(kind >= 0);
        if (qm) {
          const int lane = (int)__lane_id();
          const int leader = __ffsll((long long)qm) - 1;
          int base = 0;
          if (lane == leader) base = atomicAdd(qcount, __popcll(qm));
          base = __shfl(base, leader);
          if (kind >= 0) {
            const int slot = base + __popcll(qm & ((1ull << lane) - 1ull));
            const uint32_t idx = (uint32_t)cell * T + tr;
            if (q16) qidx16[slot] = (uint16_t)idx;
            else qidx32[slot] = idx;
            qval[slot] = kind == ITEM_POIS_ADD ? -val : val;
          }
        }
        if (kind < 0) {
          if (tn_s && mode != MODE_REPLAY) st_nt(tn_s, tr * AG + cell, n);
          if (update) {
            q[fr] -= n;
            q[to] += n;
          }
        }
      }
      {
        int4* pv = (int4*)pop;
#pragma unroll
        for (int c4 = 0; c4 < CP / 4; c4++)
          pv[c4 * AG + cell] = make_int4(q[4 * c4], q[4 * c4 + 1], q[4 * c4 + 2], q[4 * c4 + 3]);
      }
    }
    __syncthreads();
    // (D) the queued large draws (PTRS / BTRD), every lane busy.
    if (mode != MODE_REPLAY) {
      const int nq = qcount[0];
      if (nq > 0) {
        for (int i = tid; i < nq; i += nt) {
          const uint32_t idx = q16 ? (uint32_t)qidx16[i] : qidx32[i];
          const int cell = (int)(idx / T), tr = (int)(idx - (uint32_t)cell * T);
          // The kind from the value: Poisson means queued are >= the inline
          // limit, binomial probabilities lie in (0, 1), coupled Poisson
          // increments are stored negated.
          const float v = qval[i];
          const int kind = v < 0.f ? ITEM_POIS_ADD : (v < 1.f ? ITEM_BINOM : ITEM_POIS);
          int ni = mode == MODE_MBP ? itn[tr * AG + cell] : 0;
          int n = draw_item_inline(kind, fabsf(v), ni, k0, k1, c0, (uint32_t)cell, (uint32_t)tr, jb.key);
          if (tn_s) st_nt(tn_s, tr * AG + cell, n);
          if (update && n != 0) {
            int fr = 0, to = 0;
#pragma unroll
            for (int k = 0; k < T; k++)
              if (k == tr) { fr = FROM[k]; to = TO[k]; }
            atomicSub(&pop[pidx(fr, cell, AG)], n);
            atomicAdd(&pop[pidx(to, cell, AG)], n);
          }
        }
        __syncthreads();
        if (tid == 0) qcount[0] = 0;
      }
    }
    // (E) infectivity for the next step.
    for (int cell = tid; cell < AG; cell += nt) {
      const int4* pv = (const int4*)pop;
      float d = 0.f;
#pragma unroll
      for (int c4 = 0; c4 < CP / 4; c4++) {
        int4 v = pv[c4 * AG + cell];
        int pc[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const int c = 4 * c4 + k;
          if (c < C && INFC[c]) {
            float w = sinf[c];
            if (w != 0.f) d += w * (float)pc[k];
          }
        }
      }
      sdinf[cell] = d;
    }
    pool_susceptibles(pop, spool, AG, tid, nt);
    __syncthreads();
  }
  for (int cell = tid; cell < AG; cell += nt) {
#pragma unroll
    for (int c = 0; c < C; c++) ((gint*)jb.state)[c * AG + cell] = pop[pidx(c, cell, AG)];
  }
}

}  // namespace beepdev

extern "C" __global__ void __launch_bounds__(BEEP_SPEC_MAX_BLOCK, beepspec::MIN_WAVES)
    beep_fused_spec_sim(const beepdev::Job* jobs, uint64_t seed, uint32_t stream) {
  beepdev::fused_spec_body<beepdev::MODE_SIM>(jobs, seed, stream);
}
// The coupled proposal needs more registers than a simulation: at most five
// resident wavefronts per SIMD (96 registers) keep its spills small.
extern "C" __global__ void __launch_bounds__(BEEP_SPEC_MAX_BLOCK, beepspec::MIN_WAVES < 5 ? beepspec::MIN_WAVES : 5)
    beep_fused_spec_mbp(const beepdev::Job* jobs, uint64_t seed, uint32_t stream) {
  beepdev::fused_spec_body<beepdev::MODE_MBP>(jobs, seed, stream);
}
extern "C" __global__ void __launch_bounds__(BEEP_SPEC_MAX_BLOCK, beepspec::MIN_WAVES)
    beep_fused_spec_replay(const beepdev::Job* jobs, uint64_t seed, uint32_t stream) {
  beepdev::fused_spec_body<beepdev::MODE_REPLAY>(jobs, seed, stream);
}
extern "C" __global__ void __launch_bounds__(BEEP_SPEC_MAX_BLOCK, beepspec::MIN_WAVES)
    beep_fused_spec_any(const beepdev::Job* jobs, uint64_t seed, uint32_t stream) {
  beepdev::fused_spec_body<beepdev::MODE_ANY>(jobs, seed, stream);
}
