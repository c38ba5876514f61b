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
  const float* D;      // [rows][Apad] infectivity, rows = b*G + g
  const float* M;      // [Apad][Apad] off-diagonal mixing, M[a][j] (zero padded)
  float* I;            // [rows][Apad] output
  const float* diag;   // [A]
  const float* fac;    // [A]
  const Job* jobs;     // row b*G+g belongs to jobs[b] (m(t) of its slot)
  int rows, A, Apad, G, s, off_geo;
};

typedef float f16v __attribute__((ext_vector_type(16)));

// One 64x64 output tile of C = A x B (A [rows][lda], B [K][ldb], K a multiple
// of 32, tile origin (tm, tn)) per workgroup of 4 wavefronts; each wavefront
// owns a 32x32 quadrant computed with v_mfma_f32_32x32x2f32 over K in chunks
// of 32 staged through LDS, double-buffered: the global loads of chunk k+1 are
// issued before the matrix instructions of chunk k, and one barrier separates
// chunks.  The A tile is row-padded to 33 floats against bank conflicts.
// Accumulator layout of the 32x32 MFMA: lane l holds column l % 32 and, in
// register v, row 8*(v/4) + 4*(l/32) + v%4.
__device__ inline f16v mfma_tile_64x64(const float* __restrict__ A, int lda, int rows, const float* __restrict__ B,
                                       int ldb, int K, int tm, int tn) {
  __shared__ float As[2][64][33];
  __shared__ float Bs[2][32][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f16v acc;
#pragma unroll
  for (int v = 0; v < 16; v++) acc[v] = 0.f;
  // Each thread stages 8 elements of A and 8 of B per chunk.
  float ra[8], rb[8];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      int i = tid + q * 256;
      int r = i >> 5, k = i & 31;
      int gr = tm + r;
      ra[q] = gr < rows ? A[(size_t)gr * lda + k0 + k] : 0.f;
      int kb = i >> 6, c = i & 63;
      rb[q] = B[(size_t)(k0 + kb) * ldb + tn + c];
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      int i = tid + q * 256;
      As[buf][i >> 5][i & 31] = ra[q];
      Bs[buf][i >> 6][i & 63] = rb[q];
    }
  };
  load(0);
  store(0);
  __syncthreads();
  const int nk = K / 32;
  for (int kc = 0; kc < nk; kc++) {
    const int buf = kc & 1;
    if (kc + 1 < nk) load((kc + 1) * 32);
#pragma unroll
    for (int kk = 0; kk < 32; kk += 2) {
      float a = As[buf][wm + (lane & 31)][kk + (lane >> 5)];
      float b = Bs[buf][kk + (lane >> 5)][wn + (lane & 31)];
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
    }
    if (kc + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }
  return acc;
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
__global__ void __launch_bounds__(256) k_dense_contract(DenseMix p) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int tm, tn;
  if (!xcd_tile((p.rows + 63) / 64, p.Apad / 64, &tm, &tn)) return;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f16v acc = mfma_tile_64x64(p.D, p.Apad, p.rows, p.M, p.Apad, p.Apad, tm, tn);
  const int col = tn + wn + (lane & 31);
  if (col >= p.A) return;
  const float dg = p.diag[col], fc = p.fac[col];
#pragma unroll
  for (int v = 0; v < 16; v++) {
    int row = tm + wm + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
    if (row < p.rows) {
      int b = row / p.G;
      float m = p.jobs[b].tab[p.off_geo + p.s];
      float idg = dg * p.D[(size_t)row * p.Apad + col];
      p.I[(size_t)row * p.Apad + col] = (idg + acc[v]) * m + fc * idg * (1.f - m);
    }
  }
}

// Plain GEMM with the same tiling (no epilogue) for the tests of the
// accumulator layout: C[rows][Apad] = A[rows][Apad] x B[Apad][Apad].
__global__ void __launch_bounds__(256) k_dbg_gemm(const float* A_, const float* B_, float* C_, int rows, int Apad) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int tm, tn;
  if (!xcd_tile((rows + 63) / 64, Apad / 64, &tm, &tn)) return;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f16v acc = mfma_tile_64x64(A_, Apad, rows, B_, Apad, Apad, tm, tn);
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
  const float* Dr = p.D + (size_t)row * p.Apad;
  float idg = p.diag[j] * Dr[j];
  float imp = 0.f;
  for (int e = p.in_ptr[j]; e < p.in_ptr[j + 1]; e++) imp += p.in_w[e] * Dr[p.in_src[e]];
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
  ((gfloat*)D)[((size_t)blockIdx.x * G + g) * Apad + j] = d;
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
    const float foi = foi_mean(tab[tp->off_beta + s], af_row(tp, tab, s)[j], nmi, tab[tp->off_efoi + s],
                               gptr(tp->agedist)[g]);
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
    ((gfloat*)D)[((size_t)b * G + g) * Apad + j] = d;
  }
}

}  // namespace beepdev
