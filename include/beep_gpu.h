/*
 * beep_gpu.h -- C ABI of the MI355X (gfx950) implementation of BEEP's
 * stochastic simulation and likelihood hot path.
 *
 * The reference (ScottishCovidResponse/BEEP) has no FFI: its boundary is the
 * C++ class API of State / ObservationModel / Mbp (src/state.hh:11-133,
 * src/obsmodel.hh:8-39, src/mbp.hh:15-82) fed by the immutable Details /
 * Data / Model objects built from the TOML spec.  This header is the thin
 * C layer those classes' replacements call (beep_amd/csrc/host/gpu_state.hh
 * mirrors the C++ signatures on top of it).  Every entry point below names
 * the reference method whose arithmetic it carries.
 *
 * Conventions
 *  - All functions return BEEP_OK (0) or a negative error code; the message
 *    of the last error on a context is beep_last_error().
 *  - Host arrays are copied; the context owns every device buffer.
 *  - One context per host thread; calls on a context are queued on its own
 *    HIP stream and synchronise only where a result comes back to the host.
 *  - Index order of host-visible arrays is the reference's own:
 *      pop[s][area][comp][dp], transnum/transmean[s][area][trans][dp]
 *    (src/state.hh:24-28), one strain, dp == age group.
 *  - Counts are int32, rates FP32 on the device; log-likelihood sums and
 *    Metropolis-Hastings ratios are FP64.
 */
#ifndef BEEP_GPU_H
#define BEEP_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes (the reference exits through emsg/emsgEC, src/utils.cc:296-316) ---- */
#define BEEP_OK        0
#define BEEP_EINVAL   (-1)  /* bad argument or spec (emsg)                                 */
#define BEEP_ENOMEM   (-2)  /* allocation failed or over the HBM budget                    */
#define BEEP_EHIP     (-3)  /* HIP runtime error                                            */
#define BEEP_ENAN     (-4)  /* NaN transition mean (emsgEC("State",4), src/state.cc:429)    */
#define BEEP_EBOUNDS  (-5)  /* parameters outside the prior (Mbp::mbp FAIL, src/mbp.cc:254) */
#define BEEP_ELARGE   (-6)  /* transition mean > 1e9 (emsgEC("Utils",6), src/utils.cc:150)  */

/* ---- model specification: what Data / Model hold once the TOML is read ---- */

/* Prior types of Model::prior (src/model.cc:1590-1659) and Model::inbounds (2625-2669). */
#define BEEP_PRIOR_FIXED    0   /* a = value                         */
#define BEEP_PRIOR_UNIFORM  1   /* a = min, b = max                  */
#define BEEP_PRIOR_EXP      2   /* a = rate                          */
#define BEEP_PRIOR_NORMAL   3   /* a = mean, b = sd                  */
#define BEEP_PRIOR_DIRFLAT  4   /* flat Dirichlet weight, Pr = -value */
#define BEEP_PRIOR_GAMMA    5   /* a = shape, b = rate: the MDir(k) weights of a modified
                                   Dirichlet group (src/model.cc:3025-3050), gamma_probability
                                   (src/utils.cc:115-120)                                    */

typedef struct { int32_t type; double a, b; } beep_prior;

/* Compartment after Erlang expansion (src/inputs_comps.cc:33-46). */
typedef struct {
  int32_t inf_param;            /* parameter giving the infectivity; -1 = zero       */
  const int32_t* mean_param;    /* [G] parameter of the mean residence time; NULL if
                                   the compartment has no non-infection exit          */
  int32_t shape;                /* Erlang shape (transrate = shape / mean)           */
} beep_comp;

/* Transition (src/struct.hh:111-121); internal Erlang transitions first. */
typedef struct {
  int32_t from, to;
  int32_t infection;            /* 1 for the single infection transition             */
  const int32_t* prob_param;    /* [G] branching weight parameters, NULL = single exit */
} beep_trans;

/* Piecewise-linear spline (Model::create_disc_spline, src/model.cc:1020-1053), with
 * the smoothing prior of Model::spline_prior (src/model.cc:1662-1704). */
#define BEEP_SMOOTH_NONE   0
#define BEEP_SMOOTH_NORMAL 1            /* 'smooth': normal density on knot differences     */
#define BEEP_SMOOTH_LOG    2            /* 'log_smooth': log-normal density on knot ratios  */
typedef struct {
  int32_t nknot;
  const double* t;              /* knot times (days from start)                      */
  const int32_t* param;         /* [nknot] parameter at each knot                    */
  const double* multfac;        /* [nknot] multiplying factor at each knot           */
  int32_t factor_param;         /* overall factor parameter; -1 = 1                  */
  int32_t smooth_type;          /* BEEP_SMOOTH_*                                     */
  const double* smooth;         /* [nknot] smoothing sd between knot i and i+1, or null
                                   with BEEP_SMOOTH_NONE                              */
} beep_spline;

/* Dirichlet group normalised by Model::dirichlet_correct (src/model.cc:2685-2695). */
typedef struct { int32_t n; const int32_t* param; const double* frac; } beep_dirichlet;

/* Modification of the age contact matrix over time (Model::create_Ntime,
 * src/model.cc:1254-1301; age_mixing_modify / age_mixing_perturb): at step s
 * the rows and columns of the listed age groups are multiplied by the spline's
 * value.  With any, the age matrix and the NGM ratio of beta vary by step. */
typedef struct { int32_t spline; int32_t nage; const int32_t* ages; } beep_agemod;

/* Linear covariate on the area factor (Model::create_areafactor, src/model.cc:1182-1235);
 * values are centred on load as Data::read_covars does (src/data.cc:557-565). */
typedef struct { int32_t param; const double* value; /* [A][period], after any log transform */ } beep_covar;

/* Prediction-mode model modifications (Inputs 'modification', src/inputs.cc:
 * 1990-2060; their tables Model::model_modification... src/model.cc:1440-1580;
 * their use src/model.cc:1044-1049,1349-1356, src/state.cc:415,446-453). */
#define BEEP_MOD_TRANS_RATE 0   /* trans_rate_fac: every mean of the listed transitions x factor */
#define BEEP_MOD_BETA       1   /* beta_fac: beta x factor                                  */
#define BEEP_MOD_EFOI       2   /* efoi_fac: the external force of infection x factor       */
#define BEEP_MOD_SPLINE_FAC 3   /* spline_fac: the spline's value x factor                  */
#define BEEP_MOD_SPLINE_SET 4   /* spline_set: the spline's value set to factor             */
typedef struct {
  int32_t type;
  int32_t start, end;           /* steps [start, end)                                  */
  double factor;                /* the factor, or spline_set's value                    */
  int32_t ntr;   const int32_t* tr;    /* trans_rate_fac: the transitions               */
  int32_t narea; const int32_t* area;  /* geo_filt: the areas; 0 = every area          */
  int32_t ndp;   const int32_t* dp;    /* democats_filt: the dps; 0 = every dp          */
  int32_t spline;               /* spline_fac / spline_set: the spline                  */
  int32_t strain;               /* beta_fac / efoi_fac: the strain (0 with one strain)  */
} beep_mod;

/* Observation model functions (src/consts.hh ObsModelFunc). */
#define BEEP_OBS_NORMAL          0  /* sd given                                            */
#define BEEP_OBS_NORMAL_PERCENT  1  /* sd = (y(1-eps) + max_graph*eps)*percent/100          */
#define BEEP_OBS_POISSON         2
#define BEEP_OBS_NEGBIN          3  /* shape                                               */
#define BEEP_OBS_UNKNOWN   999999997.0  /* y value meaning "unknown" (contributes 0)       */
#define BEEP_OBS_THRESH    999999998.0  /* y value meaning "below threshold"               */

/* One observation (src/struct.hh:340-354): sums factor * transnum over steps
 * [sett_i, sett_f), the listed areas, dps and transitions, plus factor * pop
 * over the listed compartments (src/obsmodel.cc:74-109). */
typedef struct {
  double y, sd, shape, factor, thresh;
  double percent;               /* BEEP_OBS_NORMAL_PERCENT only                      */
  int32_t type, datatable, graph;
  int32_t sett_i, sett_f;
  int32_t narea; const int32_t* area;
  int32_t ndp;   const int32_t* dp;
  int32_t ntr;   const int32_t* tr;
  int32_t nco;   const int32_t* co;
  int32_t factor_spline;        /* 1 + the index of the obs_spline multiplying each step's
                                   count by its value at that step (src/obsmodel.cc:85-88);
                                   0 = none, so a zero-initialised observation has none;
                                   transition observations only                         */
} beep_obs;

typedef struct {
  int32_t A, G, C, T;           /* areas, age groups, compartments, transitions      */
  int32_t period;               /* days                                               */
  int32_t steps_per_day;        /* details.division_per_time (steps_per_unit_time)   */
  int32_t nparam;
  const beep_prior* prior;      /* [nparam]                                          */
  const beep_comp* comp;        /* [C]                                               */
  const beep_trans* trans;      /* [T]                                               */
  int32_t nspline;
  const beep_spline* spline;
  int32_t R_spline;             /* spline giving R(t)                                */
  int32_t Rfactor_param;        /* strain Rfactor, -1 = 1                             */
  int32_t efoi_spline;          /* external force of infection, -1 = none             */
  const double* efoi_agedist;   /* [G]                                               */
  int32_t geo_spline;           /* m(t) of geo_mixing_modify, -1 = unity              */
  const double* N_raw;          /* [nage][nage] age contact table (Data::age_mixing_matrix
                                   input), NULL = unit matrix for one age group       */
  const double* M_diag;         /* [A]   raw geographic mixing (before geo_normalise) */
  const int32_t* M_rowptr;      /* [A+1] row c lists off-diagonal destinations        */
  const int32_t* M_to;          /* [nnz]                                             */
  const double* M_val;          /* [nnz]                                             */
  const int32_t* pop_init;      /* [A][C][G] initial populations                     */
  int32_t ncovar;
  const beep_covar* covar;
  const int32_t* sus_param;     /* [G] susceptibility parameters, NULL = 1            */
  int32_t ndirichlet;
  const beep_dirichlet* dirichlet;
  int32_t nobs;
  const beep_obs* obs;
  int32_t nsection;             /* PMCMC observation sections (0 = none); section k */
  const int32_t* section_ti;    /* covers steps [section_ti[k], section_tf[k])        */
  const int32_t* section_tf;    /* (ObservationModel::split_observations, 326-348)    */
  /* Further area-factor terms (Model::create_areafactor, src/model.cc:1182-1216). */
  const int32_t* area_effect_param; /* [A] parameter multiplying area c's factor
                                   (area_effect), NULL = none                          */
  int32_t nlevel;               /* level effects (level_effect), 0 = none             */
  const int32_t* level_param;   /* [nlevel] parameter of each level                   */
  const int32_t* level_map;     /* [period][A] level of area c on day t               */
  int32_t nagemod;              /* age contact matrix modifications, 0 = none          */
  const beep_agemod* agemod;
  /* Demographic categories beyond age (democats, src/inputs.cc:642-660): the G
   * groups are then the combinations ("dp", age fastest, src/data.cc:55-73). */
  int32_t nage;                 /* age groups; N_raw is [nage][nage]                   */
  const int32_t* dp_age;        /* [G] age group of each dp; NULL = dp is the age group
                                   (nage = G)                                           */
  int32_t nsus_extra;           /* further susceptibility factors, 0 = none            */
  const int32_t* sus_extra;     /* [nsus_extra][G] parameters (-1 = 1); a dp's
                                   susceptibility is the product with sus_param's
                                   (Model::create_susceptibility, src/model.cc:1144-1178) */
  /* Strains (strains, src/inputs.cc:662-690): the last category, so dp =
   * strain * (G / nstrain) + the dp within a strain.  Each strain infects with
   * its own force of infection, beta = Rfactor * R(t) / (its NGM ratio), and
   * draws on the susceptibles of every strain's dp, pooled (src/state.cc:380-430). */
  int32_t nstrain;              /* 0 or 1 = one strain                                 */
  const int32_t* strain_Rfactor;/* [nstrain] Rfactor parameter of each strain, -1 = 1  */
  /* Prediction-mode modifications (the reference applies them in its
   * prediction mode only); 0 = none. */
  int32_t nmod;
  const beep_mod* mod;
  /* R(t) per area (R_spline entries naming an area, Model::update_R_ref,
   * src/model.cc:384-430; beta per area, Model::calculate_beta_from_R,
   * src/model.cc:1318-1347): area c follows spline rgroup_spline[g] of its
   * group g = rgroup_area[c], with beta = Rfactor * R(t) / (the NGM ratio over
   * the populations of the group's areas).  0 = R_spline for every area.  One
   * strain only. */
  int32_t nrgroup;
  const int32_t* rgroup_spline; /* [nrgroup]                                        */
  const int32_t* rgroup_area;   /* [A] group of each area                          */
  /* The external force per strain and area (efoi_spline entries naming a
   * strain or an area, Model::update_efoi_ref, src/model.cc:432-480, used at
   * src/state.cc:410-424): (strain st, area c) follows group
   * e = egroup_cell[st * A + c], efoi(t) = spline egroup_spline[e] / popsize
   * times the age weights egroup_agedist[e][dp].  0 = efoi_spline with
   * efoi_agedist everywhere. */
  int32_t negroup;
  const int32_t* egroup_spline; /* [negroup]                                        */
  const int32_t* egroup_cell;   /* [nstrain][A]                                     */
  const double* egroup_agedist; /* [negroup][G]                                     */
} beep_spec;

/* Host only (no device is touched): State::set_param's FP64 tables for one
 * parameter vector of `spec`, in the layout of beep_get_tables, plus
 * Model::prior and Model::inbounds.  Any output pointer may be NULL. */
int beep_spec_tables(const beep_spec* spec, const double* paramval, double* rate, double* sus, double* inf,
                     double* Ntime, double* beta_af, double* efoi, double* geo, double* Pr, int32_t* inbounds,
                     char* err, int errlen);

/* ---- context ---- */
typedef struct beep_ctx beep_ctx;

/* Which trajectories an instance keeps in HBM (State holds all three,
 * src/state.hh:24-28; MBPs need transnum and transmean). */
#define BEEP_KEEP_TRANSNUM   1
#define BEEP_KEEP_TRANSMEAN  2
#define BEEP_KEEP_POP        4
#define BEEP_KEEP_ALL        7
/* Option bit of `keep`: always use the step-wise path (populations in HBM,
 * one launch per kernel per step, captured in a hipGraph) even when an
 * instance fits the fused per-instance kernel.  The step-wise path is chosen
 * automatically for dense mixing and for instances too large for LDS. */
#define BEEP_STEPWISE       16
/* Option bit of `keep`: use the generic simulation kernels instead of the
 * ones specialised (compiled at context creation) for the model's
 * compartment structure.  Both give identical trajectories. */
#define BEEP_GENERIC_KERNEL 32

/* Builds Data::generate_matrices' normalised mixing (src/data_matrix.cc:14-202),
 * the observation index (ObservationModel::initialise_obs_change, obsmodel.cc:174-217)
 * and device buffers for `capacity` instances and `nslot` parameter slots.
 * hbm_budget (bytes, 0 = unlimited) bounds the allocation (BEEP_ENOMEM above it). */
int beep_create(const beep_spec* spec, int device, size_t hbm_budget, int capacity, int nslot,
                int keep, beep_ctx** out);
int beep_destroy(beep_ctx* ctx);
const char* beep_last_error(const beep_ctx* ctx);
/* Sizes: A, G, C, T, S (steps), nparam, nobs, ndatatable, capacity, nslot. */
int beep_get_dims(const beep_ctx* ctx, int32_t* dims10);
size_t beep_device_bytes(const beep_ctx* ctx);
/* The simulation path of the context: BEEP_PATH_STEPWISE (generic step-wise
 * kernels), BEEP_PATH_FUSED (generic fused kernel), BEEP_PATH_FUSED_SPEC or
 * BEEP_PATH_STEPWISE_SPEC (the same with the kernels specialised for the
 * compartment structure); negative on error. */
#define BEEP_PATH_STEPWISE   0
#define BEEP_PATH_FUSED      1
#define BEEP_PATH_FUSED_SPEC 2
#define BEEP_PATH_STEPWISE_SPEC 3
int beep_path(const beep_ctx* ctx);

/* State::set_param (src/state.cc:66-91): for each i, derives the tables of
 * parameter vector paramval[i*nparam..] (Model::dirichlet_correct,
 * create_transrate, create_disc_spline, create_susceptibility,
 * create_areafactor, create_Ntime, calculate_beta_from_R) into table slot
 * slot[i].  Pr_out[i] (may be NULL) receives Model::prior; inb_out[i] (may be
 * NULL) receives Model::inbounds (1/0).  Slots whose parameters are outside
 * the prior are still derived when the derivation is defined. */
int beep_set_param(beep_ctx* ctx, int n, const int32_t* slot, const double* paramval,
                   double* Pr_out, int32_t* inb_out);

/* Copies the FP64 tables derived for a slot (tests): rate[T][G], sus[G], inf[C],
 * Ntime[G][G] ([S][G][G] when the spec has age matrix modifications),
 * beta_af[S][A] (beta * areafactor), efoi[S][A][G] (the external force of each
 * cell: efoi(t) of its strain and area times its age weight), geo[S]. */
int beep_get_tables(beep_ctx* ctx, int slot, double* rate, double* sus, double* inf, double* Ntime,
                    double* beta_af, double* efoi, double* geo);

/* ---- hot path ---- */

/* Simulation flags.  (The reference's deterministic dynamics keep fractional
 * means as transition counts, src/state.cc:627; the int32 device state does not
 * represent them, so only stochastic dynamics run here.) */
#define BEEP_SIM_PLAIN          2   /* one launch per instance per step (the unbatched
                                       path); default is one persistent launch per batch */

/* State::simulate(ti, tf) (src/state.cc:596-638) for n instances at once:
 * instance inst[i] uses parameter slot slot[i] and Philox key key[i]; when
 * t0 == 0 the state starts from the spec's initial populations
 * (State::pop_init), otherwise from the instance's current state (pop at t0).
 * Draws are keyed by (seed, stream, key, step, area*G+dp, transition), never by
 * the instance index, so a result does not depend on what else is batched. */
int beep_simulate(beep_ctx* ctx, int n, const int32_t* inst, const int32_t* slot, const uint32_t* key,
                  int t0, int t1, uint64_t seed, uint32_t stream, int flags);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops or harnesses that forbid
 * hipMalloc): beep_simulate's batched path (same arguments, same results for
 * the same keys) that allocates and frees nothing.  Its job descriptors live
 * in a fixed 1 MiB device array of the library, one per device, so one call
 * takes at most 1 MiB / sizeof(job descriptor) instances (10082 on x86-64;
 * more is refused with BEEP_EINVAL).  Fused-path contexts only (a step-wise
 * context is refused with BEEP_EINVAL).  Every call on a device shares that
 * device's array: calls that overlap on one device (from several host
 * threads) race; each call returns synchronised with the context's stream. */
int beep_simulate_noalloc(beep_ctx* ctx, int n, const int32_t* inst, const int32_t* slot, const uint32_t* key,
                          int t0, int t1, uint64_t seed, uint32_t stream);

/* ObservationModel::calculate (src/obsmodel.cc:22-35) when section < 0, else
 * calculate_section(section) (39-58); L_out[i] = log-likelihood (FP64).
 * EF_dt_out (may be NULL) receives get_EF_datatable (309-322) as
 * [n][ndatatable] (-2 * per-table sums). */
int beep_obs_loglik(beep_ctx* ctx, int n, const int32_t* inst, int section, double* L_out,
                    double* EF_dt_out);

/* ObservationModel::get_obs_value (src/obsmodel.cc:62-70) for one instance: [nobs]. */
int beep_obs_values(beep_ctx* ctx, int inst, double* value_out);

/* State::likelihood (src/state.cc:812-826): sum over all cells of
 * log Pois(transnum | transmean), FP64. */
int beep_pois_loglik(beep_ctx* ctx, int n, const int32_t* inst, double* L_out);

/* Mbp::mbp (src/mbp.cc:252-330) for n (initial, proposed) instance pairs:
 * proposed instance prop[i] is re-simulated with parameter slot slot[i],
 * coupled to initial instance init[i]: where the proposed mean rises,
 * n_p = n_i + Pois(lp - li); where it falls, n_p = Bin(n_i, lp/li); equal
 * means keep n_i; cells with simu_mask[i][s][area] != 0 (slicetime / fixed
 * tree, src/mbp.cc:656-716) draw afresh.  simu_mask may be NULL.
 * inf_update is accepted for interface parity: INF_UPDATE (0) and
 * INF_DIF_UPDATE (1) give the same infectivity map, which the kernel
 * recomputes from the proposed populations every step. */
int beep_mbp(beep_ctx* ctx, int n, const int32_t* init, const int32_t* prop, const int32_t* slot,
             const uint32_t* key, const uint8_t* simu_mask, uint64_t seed, uint32_t stream, int inf_update);

/* Mbp::get_al + MVN::MH (src/mbp.cc:334-361, src/mvn.cc:233-249), FP64 on the
 * device: mode 0 = CUTOFF (al = exp(dlogq) if EF_p < EFcut), 1 = INVT
 * (al = exp(-invT/2 (EF_p - EF_i) + dlogq)); dlogq[i] = Pr_p - Pr_i plus any
 * proposal-density term; EF values are those of the last beep_obs_loglik on
 * each instance.  al_out[i] = min(al,1); accepted[i] = (u < al) with u a
 * Philox uniform keyed by (seed, stream, key[i]). */
int beep_accept(beep_ctx* ctx, int n, const int32_t* init, const int32_t* prop, int mode, double invT,
                double EFcut, const double* dlogq, const uint32_t* key, uint64_t seed, uint32_t stream,
                double* al_out, uint8_t* accepted);

/* ---- mixed batches across contexts ----
 * One call runs simulations and MBP proposals of many instances that may
 * belong to different contexts (different area counts, age groups, horizons):
 * the batch-submitting counterpart of the reference drivers' one-particle-at-
 * a-time loops (ABC-SMC particles, MC3 chains, src/abcsmc.cc:42-81,
 * src/mc3.cc:60-66).  Work is grouped by size class, each group sorted
 * longest first and launched as one fused kernel whose workgroups each take
 * one instance.  Results are those of the single-context calls with the same
 * keys.  Per context, the instances written (simulated, proposed) must be
 * distinct and no MBP's initial instance may be written by the same batch.
 * Contexts on the step-wise path (dense mixing, or an instance beyond LDS)
 * are refused with BEEP_EINVAL: their work goes through beep_simulate /
 * beep_mbp.  Autotuning times only size classes whose jobs all start from
 * the initial state (t0 = 0 simulations and MBPs), which a rerun recomputes
 * identically; other classes run once at the default size. */
typedef struct {
  beep_ctx* ctx;
  int32_t kind;      /* BEEP_WORK_SIMULATE or BEEP_WORK_MBP                       */
  int32_t inst;      /* simulated instance, or the proposed instance of an MBP    */
  int32_t init;      /* MBP: the initial instance                                */
  int32_t slot;      /* parameter slot                                           */
  uint32_t key;      /* Philox key                                               */
  int32_t t0, t1;    /* simulate: steps [t0, t1); MBP: whole horizon             */
} beep_work;
#define BEEP_WORK_SIMULATE   0
#define BEEP_WORK_MBP        1
/* flags of beep_run_batch */
#define BEEP_BATCH_AUTOTUNE  1   /* time the candidate workgroup sizes of each size
                                    class not yet tuned and keep the fastest (cached
                                    in memory and in the file named by the
                                    BEEP_TUNE_CACHE environment variable)           */
#define BEEP_BATCH_LOGLIK    2   /* then each item's observation log-likelihood
                                    (ObservationModel::calculate, src/obsmodel.cc:
                                    22-109, FP64) into L_out[i]; the instance's EF
                                    (-2 L) is kept for beep_accept / beep_batch_accept */
/* stats_out (may be NULL): [0] wall ms, [1] fused launches, [2] 0 (step-wise
 * contexts are not batched), [3] cell-steps (instances x areas x age groups x
 * steps). */
int beep_run_batch(int n, const beep_work* work, uint64_t seed, uint32_t stream, int flags, double* L_out,
                   double* stats_out4);
/* State::set_param (src/state.cc:66-91) for parameter slots of many contexts
 * at once, item i writing slot[i] of ctx[i] from the nparam values at
 * paramval[i]: the derivations run in parallel on the host threads, then one
 * pinned upload and one scatter kernel place every slot's tables.  Pr_out /
 * inb_out (may be NULL) as beep_set_param.  A slot may not appear twice for
 * one context; models with observation splines go through beep_set_param.
 * On a derivation error nothing is written. */
int beep_batch_set_param(int n, beep_ctx* const* ctx, const int32_t* slot, const double* const* paramval,
                         double* Pr_out, int32_t* inb_out);
/* Metropolis-Hastings acceptance (Mbp::get_al with MVN::MH, src/mbp.cc:334-361,
 * src/mvn.cc:233-249) of MBP proposals across contexts, FP64 on the device: for
 * work item i (kind BEEP_WORK_MBP), with the EF its proposed (inst) and initial
 * (init) instances last got from the observation model, al = min(1,
 * exp(-invT[i]/2 (EF_p - EF_i) + dlogq[i])) (mode 1) or exp(dlogq[i]) when
 * EF_p < EFcut (mode 0); accepted[i] = u < al with the uniform beep_accept
 * draws for the item's key. */
int beep_batch_accept(int n, const beep_work* work, int mode, const double* invT, double EFcut, const double* dlogq,
                      uint64_t seed, uint32_t stream, double* al_out, uint8_t* accepted);
/* MC3::swap_states (src/mc3.cc:153-215) for many chain ladders at once: ladder
 * l holds the chains ladder_ptr[l] .. ladder_ptr[l+1]-1 (ladder_ptr[0] = 0),
 * chain i being instance inst[i] of context ctx[i] (its EF from the last
 * observation-model evaluation) at inverse temperature invT[i], run label
 * run[i].  order_out[i]: the chain (index within its ladder) whose state chain
 * i takes; ntr/nac per chain as beep_mc3_swap.  Ladder l draws its uniforms
 * from Philox blocks (0xFFFD | stream << 16, trial, l, 0x7FFFFFFD): ladder 0
 * draws what beep_mc3_swap draws. */
int beep_batch_mc3_swap(int nladder, const int32_t* ladder_ptr, beep_ctx* const* ctx, const int32_t* inst,
                        const int32_t* run, const double* invT, uint64_t seed, uint32_t stream, int32_t* order_out,
                        int32_t* ntr_out, int32_t* nac_out);

/* ---- state transfer (State::save / load, src/state.cc:709-808; PMCMC moves) ---- */
int beep_get_transnum(beep_ctx* ctx, int inst, int32_t* out);   /* [S][A][T][G] */
int beep_get_transmean(beep_ctx* ctx, int inst, float* out);    /* [S][A][T][G] */
int beep_get_pop(beep_ctx* ctx, int inst, int32_t* out);        /* [S][A][C][G] */
int beep_get_state(beep_ctx* ctx, int inst, int32_t* pop_out);  /* current [A][C][G] */
int beep_put_state(beep_ctx* ctx, int inst, const int32_t* pop); /* current [A][C][G] */
/* Loads a transition trajectory [S][A][T][G] and rebuilds pop and transmean
 * with slot `slot` (State::initialise_from_particle, src/state.cc:294-322). */
int beep_put_transnum(beep_ctx* ctx, int inst, int slot, const int32_t* transnum);
/* Copies instance state src[i] -> dst[i] (current state and kept trajectories);
 * the device form of the particle copies of src/mpi.cc:81-165,215-327. */
int beep_copy_instances(beep_ctx* ctx, int n, const int32_t* dst, const int32_t* src);

/* Copies only the current state [C][AG] of src[i] to dst[i]. */
int beep_copy_states(beep_ctx* ctx, int n, const int32_t* dst, const int32_t* src);

/* Counterfactual slots for prediction (Simulate::model_modification,
 * src/simulate.cc:142-165): table slot dst[i] becomes slot src[i] with the
 * reproduction number multiplied by factor[i] from step from_step[i] on --
 * the `spline_fac` modification of the R spline (src/model.cc:1557-1575,
 * applied in create_disc_spline 1044-1050), which scales beta * areafactor
 * linearly.  Runs on the device; src slots must be set. */
int beep_derive_slots(beep_ctx* ctx, int n, const int32_t* dst, const int32_t* src, const double* factor,
                      const int32_t* from_step);

/* Prediction summary: out[i][s - t0] = sum over areas and dps of
 * transnum[s][area][tr][dp] of instance inst[i], s in [t0, t1). */
int beep_sum_transnum(beep_ctx* ctx, int n, const int32_t* inst, int tr, int t0, int t1, double* out);

/* ---- PMCMC particle filter ---- */
/* PMCMC::obs_prob (src/pmcmc.cc:180-207) on the device for the n particles
 * inst[] (distinct), all with parameter slot `slot` (set beforehand, as
 * obs_prob's set_param): for each observation section k in order,
 *   k > 0: particle p continues from the state of its ancestor bp_k[p]
 *          (Mpi::end_sec_swap, src/mpi.cc:215-327);
 *   simulate(section_ti[k], section_tf[k]) with draws keyed by key[p];
 *   L_p = ObservationModel::calculate_section (src/obsmodel.cc:39-58);
 *   PMCMC::bootstrap (src/pmcmc.cc:211-268): obprob += invT*Lmax +
 *     log(mean exp(invT (L - Lmax))) in FP64 and ancestors bp_{k+1} by the
 *     reference's multinomial draw and survivor / extra assignment.
 * The sections must tile [0, S) in order.  *obprob_out is the estimate;
 * backpart_out, if not null, receives [(nsection + 1)][n] ancestor rows
 * (row 0 is the identity).  The ancestry stays on the device for
 * beep_pmcmc_particle_sample.  A non-finite estimate is BEEP_ENAN
 * (src/pmcmc.cc:201).  One host synchronisation per call. */
int beep_pmcmc_obs_prob(beep_ctx* ctx, int n, const int32_t* inst, int slot, const uint32_t* key, double invT,
                        uint64_t seed, uint32_t stream, double* obprob_out, int32_t* backpart_out);

/* Mpi::particle_sample (src/mpi.cc:331-385): the trajectory (kept transnum,
 * transmean and pop, section by section) of the particle the last bootstrap
 * of the last beep_pmcmc_obs_prob call put first, traced back through the
 * ancestry, is gathered into instance dst (which may be one of the
 * particles, as particle 0 is in the reference). */
int beep_pmcmc_particle_sample(beep_ctx* ctx, int dst);

/* PMCMC::bootstrap alone (src/pmcmc.cc:211-268) on host arrays: section
 * log-likelihoods L[n] -> ancestors bp_out[n] and the section's term of the
 * estimate.  Draw j uses the same device uniform as section `sec` of
 * beep_pmcmc_obs_prob under (seed, stream). */
int beep_bootstrap(int device, int n, const double* L, double invT, uint64_t seed, uint32_t stream, int sec,
                   int32_t* bp_out, double* obprob_out);

/* ---- ABC-SMC and MC3 bookkeeping on the device ---- */
/* ABCSMC::calculate_particle_weight (src/abcsmc.cc:260-271) with
 * MVN::probability (src/mvn.cc:185-195) for a whole generation of proposals,
 * FP64:
 *   w_out[i] = exp(logprior[i]) / sum_j w_prev[j] exp(-1/2 sum_v1 sum_v2
 *              (d_v1 / size) inv_M[v1][v2] (d_v2 / size)),   d = theta[i] - prev[j],
 * over the nvar variables of the MVN kernel (the parameters that are not
 * fixed, in its order); prev [n_prev][nvar], theta [n_new][nvar],
 * inv_M [nvar][nvar]; nvar <= 32. */
int beep_abcsmc_weights(int device, int nvar, int n_prev, const double* prev, const double* w_prev, int n_new,
                        const double* theta, const double* inv_M, double size, const double* logprior,
                        double* w_out);

/* MC3::swap_states (src/mc3.cc:153-215): the n chains, in temperature order,
 * are in instances inst[] of the context with runs run[] and inverse
 * temperatures invT[]; their EF values are the instances' last ones on the
 * device (set by beep_obs_loglik over the whole horizon).  Five sweeps of
 * neighbouring swaps between chains of the same run, accepted with
 * probability exp(0.5 (invT_i - invT_i+1)(EF_i - EF_i+1)).  order_out[i] is
 * the chain whose state chain i takes (the caller relabels: chain i now uses
 * instance inst[order_out[i]]; nothing is copied); ntr_out / nac_out [n] count
 * the proposed and accepted swaps at each position. */
int beep_mc3_swap(beep_ctx* ctx, int n, const int32_t* inst, const int32_t* run, const double* invT, uint64_t seed,
                  uint32_t stream, int32_t* order_out, int32_t* ntr_out, int32_t* nac_out);

/* Waits for the context's stream. */
int beep_synchronize(beep_ctx* ctx);

/* Device time (ms, from HIP events on the context stream) of the kernels of
 * the last call of each kind: out[0] simulate / MBP / replay kernel,
 * out[1] observation log-likelihood kernel, out[2] Poisson log-likelihood
 * kernel; out[3] the number of kernel launches those times cover. */
int beep_timing(beep_ctx* ctx, double* out4);

/* ---- diagnostics (tests of the device generator and samplers) ---- */
/* Philox-4x32-10 of counters ctr[n][4] under key (k0,k1) -> out[n][4]. */
int beep_debug_philox(int device, int n, const uint32_t* ctr, uint32_t k0, uint32_t k1, uint32_t* out);
/* out[i] ~ Poisson(lam[i]) / Binomial(nn[i], p[i]) drawn with the kernels' samplers
 * from the stream of transition 0 of cell i, step 0, instance `key`. */
int beep_debug_poisson(int device, int n, const float* lam, uint64_t seed, uint32_t key, int32_t* out);
int beep_debug_binomial(int device, int n, const int32_t* nn, const double* p, uint64_t seed, uint32_t key,
                        int32_t* out);
/* C[rows][K] = A[rows][K] x B[K][K] with the dense contraction's MFMA tiling
 * (K a multiple of 64), FP32. */
int beep_debug_gemm(int device, int rows, int K, const float* A, const float* B, float* C);
/* The host-side launch checks every fused and step-wise launch passes before
 * it is enqueued (no device is touched).  kind BEEP_CHECK_FUSED, args
 * {njobs, block, lds bytes, lds limit}; kind BEEP_CHECK_STEPWISE, args {A, G,
 * Apad, n jobs, capacity}.  Returns BEEP_OK, or BEEP_EINVAL with the reason in
 * err.  On success grid_out (may be NULL) receives the grid: fused {x}, step-
 * wise {cell x, cell y, dense x, dense y, sparse x, sparse y}. */
#define BEEP_CHECK_FUSED     0
#define BEEP_CHECK_STEPWISE  1
int beep_debug_check_launch(int kind, const int64_t* args, int nargs, int64_t* grid_out, char* err, int errlen);

#ifdef __cplusplus
}
#endif
#endif /* BEEP_GPU_H */
