// Types and helpers shared by the fused simulation kernels: the generic one
// (kernels.hh, compiled with the library) and the structure-specialised one
// (fused_spec.hh, compiled at run time for one compartment structure).
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#endif

#include "device_rng.hh"

namespace beepdev {

struct TopoDev {
  int A, G, C, T, S, AG, inf_tr, spd;
  int nnz;               // off-diagonal mixing entries
  int stage_mix;         // 1: the mixing CSR is copied into LDS
  float dt;
  int off_rate, off_sus, off_inf, off_N, off_efoi, off_geo, off_beta;
  int N_tv;              // 1: the age matrix varies by step, [S][G][G] at off_N
  const int* from;       // [T]
  const int* to;         // [T]
  const float* diag;     // [A]  normalised diagonal of the area mixing
  const float* fac;      // [A]  genQ.factor
  const int* in_ptr;     // [A+1] incoming off-diagonal mixing (gather form)
  const int* in_src;     // [nnz]
  const float* in_w;     // [nnz]
  const int* mix_order;  // [A] areas by number of incoming entries, most first
  const int* pop_init;   // [C][AG]
  // Area factor [period][A]: in each parameter slot at off_af when it depends
  // on the parameters (off_af >= 0), else this context-wide table.
  const float* af;
  int off_af;
  const float* agedist;  // [nE][G] age weights of the external force of each group (efoi[S][nE] per slot)
  // Prediction-mode modifications, null when absent: multipliers of beta and
  // of the external force per step and cell, [S][AG], and of every transition
  // mean, [S][T][AG].
  const float* mod_beta;
  const float* mod_efoi;
  const float* mod_trans;
  // R(t) per area group and the external force per (strain, area) group: the
  // slot holds beta[S][nR] and efoi[S][nE]; rgrp[A] and egrp[nstrain][A] name
  // each area's and (strain, area)'s group.  nR = nE = 1: one of each.
  int nR, nE;
  const int* rgrp;
  const int* egrp;
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
  int mode;                   // MODE_SIM / MODE_MBP / MODE_REPLAY (read by MODE_ANY launches)
};

enum { MODE_SIM = 0, MODE_MBP = 1, MODE_REPLAY = 2, MODE_ANY = 3 };

// Trajectory outputs are written once and read by later kernels: global
// address space, non-temporal (streaming) stores.
typedef __attribute__((address_space(1))) int gint;
typedef __attribute__((address_space(1))) float gfloat;
// The index is 32-bit (one step's slab of an instance): with the slab base
// uniform across the wavefront the store takes a scalar 64-bit base and a
// 32-bit per-lane offset, with no 64-bit address arithmetic per lane.
__device__ inline void st_nt(int* p, uint32_t i, int v) { __builtin_nontemporal_store(v, (gint*)p + i); }
__device__ inline void st_nt(float* p, uint32_t i, float v) { __builtin_nontemporal_store(v, (gfloat*)p + i); }

// Pointers typed by address space.  A plain pointer read out of a job
// descriptor is generic, and a generic (flat) load counts against both the
// vector-memory and the LDS counters, so every later wait on LDS also waits
// for it; global and LDS reads typed as such avoid that.
typedef const __attribute__((address_space(1))) float gcfloat;
typedef const __attribute__((address_space(1))) int gcint;
typedef const __attribute__((address_space(3))) float lcfloat;
typedef const __attribute__((address_space(3))) int lcint;
__device__ inline gcfloat* gptr(const float* p) { return (gcfloat*)p; }
__device__ inline gcint* gptr(const int* p) { return (gcint*)p; }
typedef const __attribute__((address_space(1))) uint8_t gcbyte;
__device__ inline gcbyte* gptr(const uint8_t* p) { return (gcbyte*)p; }
// A structure's descriptor is read-only for every kernel: read through a
// constant-address-space pointer, its fields are uniform scalar loads.
typedef const __attribute__((address_space(4))) TopoDev ctopo;

// Area mixing (State::get_NMI, first half): I = (Idiag + Imap) m +
// fac Idiag (1 - m), with Imap = M_off^T dinf and Idiag = diag * dinf.  F and
// I are the mixing arrays' element types, qualified with the address space
// they are read from (LDS when staged).
//
// Lanes take areas in the order `order` (most incoming entries first), so
// the areas of one wavefront have similar row lengths and the gather loop,
// whose trip count is the wavefront's longest row, wastes few iterations.
// Each cell still sums its row in CSR order, four entries per iteration with
// the loads issued together: the FP32 result is the same as a plain loop's.
template <typename F, typename I>
__device__ inline void mix_cells(F* diag, F* fac, I* in_ptr, I* in_src, F* in_w, I* order, const float* sdinf,
                                 float* sI, float m, int G, int AG, int tid, int nt) {
  for (int t = tid; t < AG; t += nt) {
    const int jo = t / G, g = t - jo * G;
    const int j = order[jo], cell = j * G + g;
    float idg = diag[j] * sdinf[cell];
    float imp = 0.f;
    int e = in_ptr[j];
    const int e1 = in_ptr[j + 1];
    for (; e + 4 <= e1; e += 4) {
      const int s0 = in_src[e], s1 = in_src[e + 1], s2 = in_src[e + 2], s3 = in_src[e + 3];
      const float w0 = in_w[e], w1 = in_w[e + 1], w2 = in_w[e + 2], w3 = in_w[e + 3];
      const float d0 = sdinf[s0 * G + g], d1 = sdinf[s1 * G + g], d2 = sdinf[s2 * G + g], d3 = sdinf[s3 * G + g];
      imp += w0 * d0;
      imp += w1 * d1;
      imp += w2 * d2;
      imp += w3 * d3;
    }
    for (; e < e1; e++) imp += in_w[e] * sdinf[in_src[e] * G + g];
    sI[cell] = (idg + imp) * m + fac[j] * idg * (1.f - m);
  }
}

// The area-factor row of step s (day s / steps per day) for a slot's tables:
// the slot's own when the factor depends on the parameters, else the
// context's.  The infection mean uses beta[s] * af[day][area]
// (State::set_transmean, src/state.cc:410-415).
__device__ inline gcfloat* af_row(ctopo* tp, gcfloat* tab, int s) {
  const int day = s / tp->spd;
  gcfloat* base = tp->off_af >= 0 ? tab + tp->off_af : gptr(tp->af);
  return base + (size_t)day * tp->A;
}

// The force of infection per susceptible, beta[s] af[day][area] NMI +
// efoi[s] agedist[group] (State::set_transmean, src/state.cc:410-424), with
// its roundings fixed (two rounded products, one fused multiply-add) so that
// every kernel and compiler computes the same FP32 value.
__device__ inline float foi_mean(float beta, float af, float nmi, float efoi, float agedist) {
  const float b = beta * af;
  const float e = efoi * agedist;
  return __builtin_fmaf(b, nmi, e);
}

// Prediction-mode modifications (src/model.cc:1349-1356, src/state.cc:415,
// 446-453).  The force of infection's factors of cell `cell` at step s: af
// times beta's multiplier and agedist times the external force's; without
// modifications they stay af and agedist.
__device__ inline void mod_foi_factors(ctopo* tp, int s, int cell, float& af, float& agedist) {
  if (tp->mod_beta) af *= gptr(tp->mod_beta)[(size_t)s * tp->AG + cell];
  if (tp->mod_efoi) agedist *= gptr(tp->mod_efoi)[(size_t)s * tp->AG + cell];
}

// The group-dependent factors of the force of infection of cell (area j,
// dp g) at step s, read from the slot tables: beta of the area's R group, the
// external force of its (strain, area) group and that group's age weight
// (State::set_transmean, src/state.cc:410-424, with
// beta[st][c][s] and efoi_spl_ref[st][c]).  dpm = dps per strain.
__device__ inline void foi_groups(ctopo* tp, gcfloat* tab, int s, int j, int g, int dpm, float& beta, float& efoi,
                                  float& agedist) {
  const int nR = tp->nR, nE = tp->nE;
  const int r = nR > 1 ? gptr(tp->rgrp)[j] : 0;
  const int e = nE > 1 ? gptr(tp->egrp)[(g / dpm) * tp->A + j] : 0;
  beta = tab[tp->off_beta + (size_t)s * nR + r];
  efoi = tab[tp->off_efoi + (size_t)s * nE + e];
  agedist = gptr(tp->agedist)[(size_t)e * tp->G + g];
}

// A transition mean times its trans_rate_fac multiplier.
__device__ inline float mod_mean(ctopo* tp, int s, int tr, int cell, float lp) {
  if (tp->mod_trans) lp *= gptr(tp->mod_trans)[((size_t)s * tp->T + tr) * tp->AG + cell];
  return lp;
}

// Element (row, k) of a matrix with K columns (K a multiple of 8) stored in
// MFMA fragment order: blocks of 32 rows x 8 columns, 1 KB each, ordered by
// row block then column group; in a block, lane l = 32 h + r of a wavefront
// holds row r's columns 4h..4h+3 as one float4.  One coalesced 16-byte load
// per lane brings a whole block into registers as the operands of four
// v_mfma_f32_32x32x2f32: column 8q + 4h + i feeds MFMA 4q + i.  The step-wise
// path keeps the infectivity D in this order, and the dense mixing matrix
// transposed in it.
__host__ __device__ inline size_t frag_at(int row, int k, int K) {
  return ((((size_t)(row >> 5) * (size_t)(K >> 3) + (size_t)(k >> 3)) * 64 + (size_t)(((k >> 2) & 1) * 32 + (row & 31)))
          << 2) + (size_t)(k & 3);
}

// Queue capacity of large draws per step (items beyond it are drawn in place).
__host__ __device__ inline int fused_qcap(int AG) { return AG < 32 ? 64 : (2 * AG < 1024 ? 2 * AG : 1024); }

// Bytes of the mixing CSR when staged in LDS.
__host__ __device__ inline size_t fused_mix_bytes(int A, int nnz) { return (size_t)4 * (4 * (size_t)A + 1 + 2 * (size_t)nnz); }

// LDS bytes a job needs: populations double-buffered, infectivity, the mixed
// infectivity I, the queue of large draws, the small per-slot tables and,
// when staged, the mixing CSR.
__host__ __device__ inline size_t fused_lds_bytes(int A, int G, int C, int T, int nnz, int stage_mix) {
  int AG = A * G;
  size_t b = (size_t)4 * (2 * (size_t)C * AG + 2 * (size_t)AG + 2 * (size_t)fused_qcap(AG) + 4 + (size_t)T * G +
                          (size_t)G * G + G + C + 2 * T);
  if (stage_mix) b += fused_mix_bytes(A, nnz);
  return b;
}

#define BEEP_FUSED_MAX_BLOCK 512
#define BEEP_SPEC_MAX_BLOCK 1024        // the specialised fused kernel: up to 16 wavefronts
#define BEEP_STEP_BLOCK 256         // cells per workgroup of the step-wise transition kernels
#define BEEP_POIS_INLINE 10.0f     // means below this are drawn in place by inversion

enum { ITEM_POIS = 0, ITEM_POIS_ADD = 1, ITEM_BINOM = 2 };

// A large draw: its own Philox stream (word2 = 0x80000000 | tr << 22 | block)
// so the value does not depend on which lane draws it or when.  The
// specialised kernels call it from one place only and have it inlined.
template <bool INLINE = false>
__device__ inline __attribute__((always_inline)) int draw_item_body(int kind, float val, int ni, uint32_t k0,
                                                                    uint32_t k1, uint32_t c0, uint32_t cell,
                                                                    uint32_t tr, uint32_t key) {
  ExtraUniforms xr;
  xr.init(k0, k1, c0, cell, key);
  xr.block = 0x80000000u | (tr << 22);
  float u = xr.next();
  if (kind == ITEM_POIS) return poisson_draw(val, u, xr);
  if (kind == ITEM_POIS_ADD) return ni + poisson_draw(val, u, xr);
  return binomial_draw(ni, val, u, xr);
}
__device__ inline int draw_item(int kind, float val, int ni, uint32_t k0, uint32_t k1, uint32_t c0, uint32_t cell,
                                uint32_t tr, uint32_t key) {
  return draw_item_body(kind, val, ni, k0, k1, c0, cell, tr, key);
}
__device__ inline __attribute__((always_inline)) int draw_item_inline(int kind, float val, int ni, uint32_t k0,
                                                                      uint32_t k1, uint32_t c0, uint32_t cell,
                                                                      uint32_t tr, uint32_t key) {
  return draw_item_body<true>(kind, val, ni, k0, k1, c0, cell, tr, key);
}

// The specialised kernel's queue of large draws names an item by
// cell * T + tr, in 16 bits when every such index fits.
__host__ __device__ inline bool spec_queue16(int T, int AG) { return (size_t)T * AG <= 65536; }

// LDS bytes of the structure-specialised kernel: populations [AG][CP] (one
// buffer, CP = C rounded up to 4), infectivity, mixed infectivity, the queue
// of large draws (room for every (cell, transition): a float value each, and
// an index of 2 or 4 bytes), the per-slot tables and, when staged, the
// mixing CSR.
// With strains, the susceptibles pooled over strains take one int per cell
// more, after the queue's indices.
// With R or external-force groups (grouped), each step's beta of every area
// and external force of every (strain, area) are staged too, with each
// (strain, area)'s group, and the age weights of every external-force group
// (nE of them).
__host__ __device__ inline size_t spec_lds_bytes(int A, int G, int C, int T, int nnz, int stage_mix, int nstrain = 1,
                                                 int grouped = 0, int nE = 1) {
  int AG = A * G, CP = (C + 3) / 4 * 4;
  size_t q = (size_t)T * AG;
  size_t b = (size_t)4 * ((size_t)AG * CP + 2 * (size_t)AG + q + 4 + (size_t)T * G + (size_t)G * G + G + C + A +
                          (size_t)nE * G + (grouped ? (size_t)A * (1 + 2 * (size_t)nstrain) : 0));
  b += (q * (spec_queue16(T, AG) ? 2 : 4) + 3) / 4 * 4;
  if (stage_mix) b += fused_mix_bytes(A, nnz);
  if (nstrain > 1) b += (size_t)4 * AG;
  return b;
}

}  // namespace beepdev
