// Host-side model: the normalised topology and the per-parameter derived
// tables that feed the device kernels.  This is the C++ host's counterpart of
// the reference's Data / Model set-up (src/data_matrix.cc, src/data.cc) and
// of State::set_param (src/state.cc:66-91), written for flat arrays.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "beep_gpu.h"

namespace beephost {

struct Compartment { int inf_param; std::vector<int> mean_param; int shape; std::vector<int> exits; };
struct Transition { int from, to, infection; std::vector<int> prob_param; };
struct Spline { std::vector<double> t, multfac, smooth; std::vector<int> param; int factor_param, smooth_type; };
struct Dirichlet { std::vector<int> param; std::vector<double> frac; };
struct Covar { int param; std::vector<double> value; /* [A][period], centred */ };

// One observation flattened to element indices into an instance's device
// trajectories: bit 31 clear -> transnum[s][tr][cell] (index s*T*AG + tr*AG + cell),
// bit 31 set -> pop[s][co][cell] (index s*C*AG + co*AG + cell).
struct Obs {
  double y, sd, shape, factor, thresh;
  int type, datatable, sett_i, sett_f;
  int ospl = -1;                // index into obs_splines, -1 = none
  std::vector<uint32_t> elem;
};

// Tables State::set_param derives from one parameter vector (FP64).
struct Tables {
  std::vector<double> rate;     // [T][G]
  std::vector<double> sus;      // [G]
  std::vector<double> inf;      // [C]
  std::vector<double> Ntime;    // [G][G], or [S][G][G] with age matrix modifications
  std::vector<double> beta;     // [S][nR] Rfactor * R(t_s) / NGM ratio of each R group (strains: R(t_s))
  std::vector<double> af;       // [period][A] area factor; empty when it does not depend
                                // on the parameters (HostModel::af_const holds it)
  std::vector<double> efoi;     // [S][nE] efoi(t_s) / popsize of each external-force group
  std::vector<double> geo;      // [S]    m(t_s)
  std::vector<double> obs_spl;  // [nobs_spline][S] observation spline values
};

class HostModel {
 public:
  // Returns an empty string on success, else the error message.
  std::string build(const beep_spec& spec);

  int A = 0, G = 0, C = 0, T = 0, S = 0, AG = 0, period = 0, spd = 0, inf_tr = -1, nparam = 0;
  double dt = 1.0;
  double popsize = 0;

  std::vector<Compartment> comp;
  std::vector<Transition> trans;
  std::vector<beep_prior> prior;
  std::vector<Spline> spline;
  std::vector<Dirichlet> dirichlet;
  std::vector<Covar> covar;
  std::vector<int> area_effect_param;                         // [A], empty = none
  std::vector<int> level_param, level_map;                    // [nlevel], [period][A]
  struct AgeMod { int spline; std::vector<int> ages; };
  std::vector<AgeMod> agemod;                                 // empty: N constant in time
  std::vector<int> sus_param;
  std::vector<std::vector<int>> sus_extra;                   // further factors, [G] each
  int nage = 0;
  std::vector<int> dp_age;                                    // [G] age group of each dp
  int nstrain = 1;                                            // strains: G = nstrain x dps per strain
  std::vector<int> strain_Rfactor;                            // [nstrain] parameter, -1 = 1
  int R_spline = -1, Rfactor_param = -1, efoi_spline = -1, geo_spline = -1;
  std::vector<double> efoi_agedist;
  std::vector<double> Rdist;          // [G] population fraction of every area's dps
  // R(t) per area group and the external force per (strain, area) group
  // (Model::update_R_ref / update_efoi_ref, src/model.cc:384-480): one group
  // each unless the spec names more.
  int nR = 1, nE = 1;
  std::vector<int> rgrp_spline, rgrp_area;                    // [nR], [A]
  std::vector<std::vector<double>> rgrp_dist;                 // [nR][G] dp fractions over the group's areas
  std::vector<int> egrp_spline, egrp_cell;                    // [nE], [nstrain][A]
  std::vector<double> egrp_agedist;                           // [nE][G]

  // Data::geo_normalise (src/data_matrix.cc:55-96), reference orientation ...
  std::vector<double> M_diag, M_val, M_factor;
  std::vector<int> M_rowptr, M_to;
  // ... and its transpose, the gather form the kernels use: Imap[j] = sum over
  // incoming (src, w) of w * dinf[src].
  std::vector<int> in_ptr, in_src;
  std::vector<double> in_w;
  std::vector<double> N_age;          // [nage][nage] normalised age mixing
  std::vector<double> N;              // [G][G] the same for each pair of dps
  std::vector<int> pop_init;          // [A][C][G]
  std::vector<double> area_pop;       // [A]
  std::vector<int> inf_state;         // infected compartments (Model::set_infected_uninfected)

  std::vector<Obs> obs;
  std::vector<int> obs_splines;       // the splines observations use, in first-use order
  int ndatatable = 1;
  std::vector<int> section_ti, section_tf;
  std::vector<std::vector<int>> section_obs;

  // The area factor depends on the parameters (area effects, covariates,
  // level effects); otherwise it is computed once, into af_const [period][A].
  bool af_param = false;
  std::vector<double> af_const;

  // Prediction-mode modifications (src/model.cc:1440-1580), as multipliers per
  // step and cell (dp of an area): of beta (beta_fac), of the external force
  // (efoi_fac) and of every transition mean (trans_rate_fac), each empty when
  // no modification of its kind is given; per spline and step, the factor of
  // spline_fac and the value of spline_set (NaN = not set), empty = none.
  std::vector<double> mod_beta;                               // [S][AG]
  std::vector<double> mod_efoi;                               // [S][AG]
  std::vector<double> mod_trans;                              // [S][T][AG]
  std::vector<std::vector<double>> spline_mult, spline_set;   // [nspline][S]

  std::vector<double> dirichlet_correct(const double* paramv) const;
  // Returns 0 or a BEEP_* error code.
  int derive(const double* paramv, Tables& out, std::string& err) const;
  // beta[s][group of area] * areafactor[day of s][area], [S][A] (FP64).
  std::vector<double> beta_af(const Tables& tb) const;
  // The external force of every cell, efoi[s][its group] * the group's age
  // weight, [S][A][G] (FP64).
  std::vector<double> efoi_cells(const Tables& tb) const;
  double prior_logpdf(const double* paramv) const;
  bool inbounds(const double* paramv) const;

 private:
  std::string build_modifications(const beep_spec& s);
  std::string build_groups(const beep_spec& s);
  std::vector<double> disc_spline(int sp, const std::vector<double>& pdir) const;
  std::vector<double> area_factor(const std::vector<double>& pdir) const;
  int strain_beta(const std::vector<double>& pd, Tables& tb, std::string& err) const;
};

}  // namespace beephost
