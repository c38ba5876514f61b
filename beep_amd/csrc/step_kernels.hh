// Step-wise path: one launch per kernel per step, state in HBM.  Used for
// instances whose populations do not fit in LDS and for dense area mixing,
// where the area-mixing contraction of a whole batch is one GEMM on the
// matrix cores:
//   Imap[(b,g)][j] = sum_a dinf[(b,g)][a] * M[a][j]       (FP32 in, FP32 acc)
// with the FOI epilogue I = (Idiag + Imap) m + fac Idiag (1 - m) fused into
// the GEMM's store (State::set_I_from_pop / update_I_from_transnum and the
// first half of State::get_NMI, src/state.cc:178-290,547-578).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hh"

namespace beepdev {

// Dense contraction operands of one launch.
struct DenseMix {
  const float* D;      // infectivity, rows b*G + g (rows padded to 64), Apad columns, fragment order
  const float* Mt;     // [Apad][Apad] off-diagonal mixing transposed, Mt[j][a] = M[a][j], fragment order
  float* I;            // [rows][Apad] output, row-major
  const float* diag;   // [A]
  const float* fac;    // [A]
  const Job* jobs;     // row b*G+g belongs to jobs[b] (m(t) of its slot)
  int rows, A, Apad, G, s, off_geo;
};

typedef float f16v __attribute__((ext_vector_type(16)));

// One 64x64 output tile of C = A x B (origin (tm, tn)) per workgroup of four
// wavefronts.  A (rows padded to a multiple of 64) and B transposed are in
// fragment order (frag_at), K a multiple of 64.  K is split four ways: each
// wavefront computes the whole 64x64 tile over its quarter of K with four
// accumulators (each 16-byte operand load feeds two MFMAs, and the four
// accumulation chains interleave), streaming its row blocks of A and column
// blocks of B from L2 straight into registers, the next block's four loads in
// flight while the current block's 16 MFMAs run.  The four partial tiles are
// then summed through LDS (48 KB): wavefront w returns the sum of quadrant w
// (rows 32 (w / 2), columns 32 (w % 2) of the tile).  2880 wavefronts for the
// 9216 x 320 x 320 product of 1024 particles of 9 groups, 2.8 per SIMD.
// Accumulator layout of the 32x32 MFMA: lane l holds column l % 32 and, in
// register v, row 8*(v/4) + 4*(l/32) + v%4.
__device__ inline f16v mfma_tile_ksplit(const float* __restrict__ Af, const float* __restrict__ Btf, int K, int tm,
                                        int tn) {
  __shared__ float red[4][3][16][64];       // [writer][its other quadrants][register][lane]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nq = K >> 3, nqw = nq >> 2, q0 = wave * nqw, qe = q0 + nqw;
  const float4* __restrict__ a4[2];
  const float4* __restrict__ b4[2];
#pragma unroll
  for (int i = 0; i < 2; i++) {
    a4[i] = (const float4*)Af + (size_t)((tm >> 5) + i) * nq * 64 + lane;
    b4[i] = (const float4*)Btf + (size_t)((tn >> 5) + i) * nq * 64 + lane;
  }
  f16v acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int v = 0; v < 16; v++) acc[i][j][v] = 0.f;
  float4 a0[2], b0[2], a1[2], b1[2];
  // A scheduling barrier after each group of loads keeps the compiler from
  // sinking them next to their MFMAs (which would wait out every load's
  // latency).  The loop body has no condition around its loads, so each
  // wait before a block's MFMAs leaves the next block's loads in flight.
  auto load = [&](float4(&a)[2], float4(&b)[2], int q) {
#pragma unroll
    for (int i = 0; i < 2; i++) {
      a[i] = a4[i][q * 64];
      b[i] = b4[i][q * 64];
    }
    __builtin_amdgcn_sched_barrier(0);
  };
  auto mfma4 = [&](float x0, float x1, float y0, float y1) {
    acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(x0, y0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(x0, y1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(x1, y0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(x1, y1, acc[1][1], 0, 0, 0);
  };
  // Block component i is the operand of the block's MFMA i.
  auto compute = [&](const float4(&a)[2], const float4(&b)[2]) {
    mfma4(a[0].x, a[1].x, b[0].x, b[1].x);
    mfma4(a[0].y, a[1].y, b[0].y, b[1].y);
    mfma4(a[0].z, a[1].z, b[0].z, b[1].z);
    mfma4(a[0].w, a[1].w, b[0].w, b[1].w);
  };
  load(a0, b0, q0);
  int q = q0;
  for (; q + 2 < qe; q += 2) {
    load(a1, b1, q + 1);
    compute(a0, b0);
    load(a0, b0, q + 2);
    compute(a1, b1);
  }
  load(a1, b1, q + 1);
  compute(a0, b0);
  compute(a1, b1);
  // Quadrant qd of writer w goes to slot qd - (qd > w) of its three.
#pragma unroll
  for (int qd = 0; qd < 4; qd++) {
    if (qd == wave) continue;
#pragma unroll
    for (int v = 0; v < 16; v++) red[wave][qd - (qd > wave)][v][lane] = acc[qd >> 1][qd & 1][v];
  }
  __syncthreads();
  f16v s;
#pragma unroll
  for (int qd = 0; qd < 4; qd++)
    if (qd == wave) s = acc[qd >> 1][qd & 1];
#pragma unroll
  for (int w2 = 0; w2 < 4; w2++) {
    if (w2 == wave) continue;
    const int slot = wave - (wave > w2);
#pragma unroll
    for (int v = 0; v < 16; v++) s[v] += red[w2][slot][v][lane];
  }
  return s;
}

// Tiles of a 64x64-tiled product for a one-dimensional grid of
// 8 * ceil(tiles / 8) workgroups.  Workgroups are dealt to the eight XCDs in
// turn (workgroup w runs on XCD w % 8), so XCD x is given the contiguous run
// of tiles [x * per, (x + 1) * per) in row-major tile order: the column tiles
// of one row tile share an XCD and its L2, which then reads that row tile of
// the left operand once instead of once per column tile.  Returns false for
// the few workgroups past the last tile (the whole workgroup returns).
__device__ inline bool xcd_tile(int ntm, int ntn, int* tm, int* tn) {
  const int total = ntm * ntn;
  const int per = (total + 7) >> 3;
  const int t = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  if (t >= total) return false;
  *tm = (t / ntn) * 64;
  *tn = (t - (t / ntn) * ntn) * 64;
  return true;
}

// The dense contraction with its FOI epilogue, one-dimensional grid of
// 8 * ceil(ceil(rows / 64) * (Apad / 64) / 8) workgroups (xcd_tile).
// Three workgroups per CU (48 KB of LDS each) hold the grid of 720
// workgroups of the 9216-row product at once: registers are capped for three
// wavefronts per SIMD.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 4))) k_dense_contract(DenseMix p) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int tm, tn;
  if (!xcd_tile((p.rows + 63) / 64, p.Apad / 64, &tm, &tn)) return;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f16v acc = mfma_tile_ksplit(p.D, p.Mt, p.Apad, tm, tn);
  const int col = tn + wn + (lane & 31);
  if (col >= p.A) return;
  const float dg = p.diag[col], fc = p.fac[col];
  // Every load of the epilogue is issued before the first store: the stores
  // to I may alias the job table and D as far as the compiler knows, so
  // loads interleaved with them would wait out one latency per row.
  const int rbase = tm + wm + 4 * (lane >> 5);
  float mv[16], dv[16];
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int row = min(rbase + 8 * (v >> 2) + (v & 3), p.rows - 1);
    mv[v] = p.jobs[row / p.G].tab[p.off_geo + p.s];
    dv[v] = p.D[frag_at(row, col, p.Apad)];
  }
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int row = rbase + 8 * (v >> 2) + (v & 3);
    if (row < p.rows) {
      const float idg = dg * dv[v];
      p.I[(size_t)row * p.Apad + col] = (idg + acc[v]) * mv[v] + fc * idg * (1.f - mv[v]);
    }
  }
}

// Plain GEMM with the same tiling (no epilogue) for the tests of the operand
// order and accumulator layout: C[rows][Apad] = A x B, A and B transposed in
// fragment order.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 4)))
k_dbg_gemm(const float* Af, const float* Btf, float* C_, int rows, int Apad) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int tm, tn;
  if (!xcd_tile((rows + 63) / 64, Apad / 64, &tm, &tn)) return;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f16v acc = mfma_tile_ksplit(Af, Btf, Apad, tm, tn);
  const int col = tn + wn + (lane & 31);
#pragma unroll
  for (int v = 0; v < 16; v++) {
    int row = tm + wm + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
    if (row < rows) C_[(size_t)row * Apad + col] = acc[v];
  }
}

// Sparse contraction for the step-wise path: one thread per (row, area j),
// Imap = sum over incoming edges of w * dinf[src], with the same epilogue.
struct SparseMix {
  const float* D;
  float* I;
  const float* diag;
  const float* fac;
  const int* in_ptr;
  const int* in_src;
  const float* in_w;
  const Job* jobs;
  int rows, A, Apad, G, s, off_geo;
};

__global__ void __launch_bounds__(256) k_sparse_contract(SparseMix p) {
  int j = blockIdx.y * blockDim.x + threadIdx.x;
  int row = blockIdx.x;
  if (j >= p.A) return;
  float idg = p.diag[j] * p.D[frag_at(row, j, p.Apad)];
  float imp = 0.f;
  for (int e = p.in_ptr[j]; e < p.in_ptr[j + 1]; e++) imp += p.in_w[e] * p.D[frag_at(row, p.in_src[e], p.Apad)];
  float m = p.jobs[row / p.G].tab[p.off_geo + p.s];
  p.I[(size_t)row * p.Apad + j] = (idg + imp) * m + p.fac[j] * idg * (1.f - m);
}

// Start of a step-wise run: the state at t0 (the spec's initial populations
// when t0 == 0) and its infectivity D[(b,g)][j] = sum_c inf_c pop_c.
// Grid (nb, ceil(AG/256)): job on x, block of cells on y.
__global__ void __launch_bounds__(256) k_step_init(const Job* __restrict__ jobs, float* __restrict__ D, int Apad) {
  const Job jb = jobs[blockIdx.x];
  ctopo* tp = (ctopo*)jb.topo;
  const int AG = tp->AG, C = tp->C, G = tp->G;
  int cell = blockIdx.y * blockDim.x + threadIdx.x;
  if (cell >= AG) return;
  gint* state = (gint*)jb.state;
  float d = 0.f;
  for (int c = 0; c < C; c++) {
    int v = jb.from_init ? gptr(tp->pop_init)[c * AG + cell] : state[c * AG + cell];
    if (jb.from_init) state[c * AG + cell] = v;
    float w = gptr(jb.tab)[tp->off_inf + c];
    if (w != 0.f) d += w * (float)v;
  }
  int j = cell / G, g = cell - j * G;
  ((gfloat*)D)[frag_at(blockIdx.x * G + g, j, Apad)] = d;
}

#define BEEP_STEP_QCAP 512

// One step s of the tau-leap for a batch, one thread per cell, grid
// (nb, ceil(AG/256)): means from the mixed infectivity I of the contraction,
// small draws in place, large ones through an LDS queue, populations updated
// in HBM, and the next step's infectivity D written.
template <int MODE>
__global__ void __launch_bounds__(BEEP_STEP_BLOCK) k_step_trans(const Job* __restrict__ jobs, const float* __restrict__ Ibuf,
                                                               float* __restrict__ D, int Apad, int s,
                                                               const uint64_t* __restrict__ rngp) {
  const uint64_t seed = rngp[0];
  const uint32_t stream = (uint32_t)rngp[1];
  const int b = blockIdx.x;
  const Job jb = jobs[b];
  ctopo* __restrict__ tp = (ctopo*)jb.topo;
  const int A = tp->A, G = tp->G, C = tp->C, T = tp->T, S = tp->S, AG = tp->AG, inf_tr = tp->inf_tr;
  gcint* tfrom = gptr(tp->from);
  gcint* tto = gptr(tp->to);
  const float dt = tp->dt;
  const int tid = threadIdx.x;
  const int cell = blockIdx.y * BEEP_STEP_BLOCK + tid;
  const bool valid = cell < AG && s >= jb.t0 && s < jb.t1;

  __shared__ int spop0[16][BEEP_STEP_BLOCK];     // populations at step s of this workgroup's cells, C <= 16
  __shared__ int spop[16][BEEP_STEP_BLOCK];      // populations after the step's transitions
  __shared__ uint32_t qcode[BEEP_STEP_QCAP];
  __shared__ float qval[BEEP_STEP_QCAP];
  __shared__ int qcount;
  if (tid == 0) qcount = 0;
  __syncthreads();

  gcfloat* tab = gptr(jb.tab);
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
    for (int c = 0; c < C; c++) {
      int v = state[c * AG + cell];
      spop0[c][tid] = v;
      spop[c][tid] = v;
      if (pt_s) st_nt(pt_s, c * AG + cell, v);
    }
    gcfloat* Ib = gptr(Ibuf) + (size_t)b * G * Apad;
    float nmi = 0.f;
    const size_t offN = tp->off_N + (tp->N_tv ? (size_t)s * G * G : 0);   // a time-varying age matrix: step s's
    for (int gg = 0; gg < G; gg++) nmi += tab[offN + g * G + gg] * Ib[(size_t)gg * Apad + j];
    if (nmi < 0.f) nmi = 0.f;
    float afc = af_row(tp, tab, s)[j], agc = gptr(tp->agedist)[g];
    mod_foi_factors(tp, s, cell, afc, agc);
    const float foi = foi_mean(tab[tp->off_beta + s], afc, nmi, tab[tp->off_efoi + s], agc);
    const bool simu = MODE == MODE_MBP ? (jb.simu ? gptr(jb.simu)[(size_t)s * A + j] != 0 : false) : true;
    CellRng rng;
    rng.init(k0, k1, c0, (uint32_t)cell, jb.key);
    if (MODE != MODE_REPLAY) rng.prime();
#pragma unroll 1
    for (int tr = 0; tr < T; tr++) {
      const int fr = tfrom[tr];
      float popu = (float)spop0[fr][tid];
      float lp = 0.f;
      if (popu > 0.f)
        lp = tr == inf_tr ? dt * popu * tab[tp->off_sus + g] * foi : dt * popu * tab[tp->off_rate + tr * G + g];
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
            if ((float)ni * pp < 10.f) n = binomial_draw(ni, pr, rng.next(), rng);
            else { kind = ITEM_BINOM; val = pr; }
          }
        }
      }
      if (kind >= 0) {
        int slot = atomicAdd(&qcount, 1);
        if (slot < BEEP_STEP_QCAP) {
          qcode[slot] = (uint32_t)tid | ((uint32_t)tr << 20) | ((uint32_t)kind << 28);
          qval[slot] = val;
          continue;
        }
        int ni = MODE == MODE_MBP ? itn[tr * AG + cell] : 0;
        n = draw_item(kind, val, ni, k0, k1, c0, (uint32_t)cell, (uint32_t)tr, jb.key);
      }
      if (tn_s && MODE != MODE_REPLAY) st_nt(tn_s, tr * AG + cell, n);
      if (update && n != 0) {
        atomicSub(&spop[fr][tid], n);
        atomicAdd(&spop[tto[tr]][tid], n);
      }
    }
  }
  __syncthreads();
  int nq = qcount < BEEP_STEP_QCAP ? qcount : BEEP_STEP_QCAP;
  for (int i = tid; i < nq; i += BEEP_STEP_BLOCK) {
    uint32_t code = qcode[i];
    int t2 = (int)(code & 0xFFFFFu), tr = (int)((code >> 20) & 0xFFu), kind = (int)(code >> 28);
    int c2 = blockIdx.y * BEEP_STEP_BLOCK + t2;
    int ni = MODE == MODE_MBP ? itn[tr * AG + c2] : 0;
    int n = draw_item(kind, qval[i], ni, k0, k1, c0, (uint32_t)c2, (uint32_t)tr, jb.key);
    if (tn_s) st_nt(tn_s, tr * AG + c2, n);
    if (update && n != 0) {
      atomicSub(&spop[tfrom[tr]][t2], n);
      atomicAdd(&spop[tto[tr]][t2], n);
    }
  }
  __syncthreads();
  if (valid) {
    float d = 0.f;
    for (int c = 0; c < C; c++) {
      int v = spop[c][tid];
      state[c * AG + cell] = v;
      float w = tab[tp->off_inf + c];
      if (w != 0.f) d += w * (float)v;
    }
    ((gfloat*)D)[frag_at(b * G + g, j, Apad)] = d;
  }
}

}  // namespace beepdev
