// Host-side model set-up and State::set_param derivations (FP64).
#include "host_model.hh"

#include <cmath>
#include <cstring>

namespace beephost {

namespace {
const double LARGE = 1e9;
const double EPSILON_FACTOR = 0.01;   // src/consts.hh:178
const double VTINY = 1e-15;           // src/consts.hh:171

// Gauss-Jordan inverse with partial pivoting.  Returns false when singular.
bool invert(std::vector<double> M, int n, std::vector<double>& inv) {
  inv.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) inv[(size_t)i * n + i] = 1.0;
  for (int col = 0; col < n; col++) {
    int piv = col;
    for (int r = col + 1; r < n; r++)
      if (std::fabs(M[(size_t)r * n + col]) > std::fabs(M[(size_t)piv * n + col])) piv = r;
    if (M[(size_t)piv * n + col] == 0) return false;
    if (piv != col)
      for (int k = 0; k < n; k++) {
        std::swap(M[(size_t)piv * n + k], M[(size_t)col * n + k]);
        std::swap(inv[(size_t)piv * n + k], inv[(size_t)col * n + k]);
      }
    double d = M[(size_t)col * n + col];
    for (int k = 0; k < n; k++) { M[(size_t)col * n + k] /= d; inv[(size_t)col * n + k] /= d; }
    for (int r = 0; r < n; r++) {
      if (r == col) continue;
      double f = M[(size_t)r * n + col];
      if (f == 0) continue;
      for (int k = 0; k < n; k++) {
        M[(size_t)r * n + k] -= f * M[(size_t)col * n + k];
        inv[(size_t)r * n + k] -= f * inv[(size_t)col * n + k];
      }
    }
  }
  return true;
}

// Dominant eigenvalue by power iteration with sum normalisation and the
// convergence test of the reference's largest_eigenvalue (src/matrix.cc:1139-1219):
// eigenvalue and normalised vector both within 1e-13 (relaxed x10 after 100 and
// again after 1000 iterations).  Returns NaN when it does not converge.
double dominant_eigenvalue(const std::vector<double>& M, int n) {
  std::vector<double> v(n, 1.0 / n), v2(n);
  double ev = 0;
  for (int loop = 0; loop <= 10000; loop++) {
    for (int i = 0; i < n; i++) {
      double s = 0;
      for (int k = 0; k < n; k++) s += M[(size_t)i * n + k] * v[k];
      v2[i] = s;
    }
    double sum = 0;
    for (int i = 0; i < n; i++) sum += v2[i];
    double ev_new = sum;
    for (int i = 0; i < n; i++) v2[i] /= sum;
    if (loop % 10 == 0) {
      double lim = 100 * VTINY;
      if (loop > 100) lim *= 10;
      if (loop > 1000) lim *= 10;
      if (ev_new - ev > -lim && ev_new - ev < lim) {
        int k = 0;
        for (; k < n; k++) {
          double d = v2[k] - v[k];
          if (d > lim || d < -lim) break;
        }
        if (k == n) return ev_new;
      }
    }
    v = v2;
    ev = ev_new;
  }
  return NAN;
}
}  // namespace

std::string HostModel::build(const beep_spec& s) {
  A = s.A; G = s.G; C = s.C; T = s.T;
  period = s.period; spd = s.steps_per_day;
  if (A <= 0 || G <= 0 || C <= 0 || T <= 0 || period <= 0 || spd <= 0) return "bad dimensions";
  if (T > 255) return "too many transitions";
  S = period * spd;
  if (S > 65533) return "too many steps (at most 65533)";
  AG = A * G;
  dt = (double)period / S;                                   // src/state.cc:384
  nparam = s.nparam;
  prior.assign(s.prior, s.prior + nparam);
  for (const beep_prior& p : prior) {
    if (p.type < BEEP_PRIOR_FIXED || p.type > BEEP_PRIOR_GAMMA) return "unknown prior type";
    if (p.type == BEEP_PRIOR_GAMMA && (p.a <= 0 || p.b <= 0)) return "gamma prior needs a positive shape and rate";
  }

  comp.resize(C);
  for (int c = 0; c < C; c++) {
    comp[c].inf_param = s.comp[c].inf_param;
    comp[c].shape = s.comp[c].shape > 0 ? s.comp[c].shape : 1;
    if (s.comp[c].mean_param) comp[c].mean_param.assign(s.comp[c].mean_param, s.comp[c].mean_param + G);
    if (comp[c].inf_param >= nparam) return "infectivity parameter out of range";
  }
  trans.resize(T);
  inf_tr = -1;
  for (int t = 0; t < T; t++) {
    const beep_trans& x = s.trans[t];
    if (x.from < 0 || x.from >= C || x.to < 0 || x.to >= C) return "transition compartment out of range";
    trans[t].from = x.from; trans[t].to = x.to; trans[t].infection = x.infection;
    if (x.prob_param) trans[t].prob_param.assign(x.prob_param, x.prob_param + G);
    if (x.infection) {
      if (inf_tr >= 0) return "more than one infection transition";
      inf_tr = t;
    }
    comp[x.from].exits.push_back(t);
  }
  if (inf_tr < 0) return "no infection transition";           // src/model.cc:186-198
  for (int t = 0; t < T; t++) {
    if (trans[t].infection) continue;
    const Compartment& c = comp[trans[t].from];
    if (c.mean_param.empty()) return "compartment with an exit has no mean parameter";
    if (c.exits.size() > 1 && trans[t].prob_param.empty()) return "branching transition has no probability parameter";
  }

  spline.resize(s.nspline);
  for (int k = 0; k < s.nspline; k++) {
    const beep_spline& x = s.spline[k];
    if (x.nknot < 2) return "spline needs two knots";
    spline[k].t.assign(x.t, x.t + x.nknot);
    spline[k].param.assign(x.param, x.param + x.nknot);
    spline[k].multfac.assign(x.multfac, x.multfac + x.nknot);
    spline[k].factor_param = x.factor_param;
    spline[k].smooth_type = x.smooth_type;
    if (x.smooth_type != BEEP_SMOOTH_NONE) {
      if (x.smooth_type != BEEP_SMOOTH_NORMAL && x.smooth_type != BEEP_SMOOTH_LOG) return "unknown spline smoothing type";
      if (!x.smooth) return "a smoothed spline needs its smoothing values";
      spline[k].smooth.assign(x.smooth, x.smooth + x.nknot);
    }
  }
  R_spline = s.R_spline; Rfactor_param = s.Rfactor_param;
  efoi_spline = s.efoi_spline; geo_spline = s.geo_spline;
  if (R_spline < 0 || R_spline >= s.nspline) return "R spline not set";
  efoi_agedist.assign(G, 0.0);
  if (s.efoi_agedist) efoi_agedist.assign(s.efoi_agedist, s.efoi_agedist + G);
  if (s.sus_param) sus_param.assign(s.sus_param, s.sus_param + G);
  sus_extra.clear();
  for (int k = 0; k < s.nsus_extra; k++) {
    sus_extra.emplace_back(s.sus_extra + (size_t)k * G, s.sus_extra + (size_t)(k + 1) * G);
    for (int th : sus_extra.back())
      if (th >= nparam) return "susceptibility parameter out of range";
  }
  nstrain = s.nstrain > 1 ? s.nstrain : 1;
  strain_Rfactor.assign(nstrain, -1);
  if (nstrain > 1) {
    if (G % nstrain) return "the groups are not a whole number per strain";
    if (!s.strain_Rfactor) return "strains need their Rfactor parameters";
    strain_Rfactor.assign(s.strain_Rfactor, s.strain_Rfactor + nstrain);
    for (int th : strain_Rfactor)
      if (th >= nparam) return "strain Rfactor parameter out of range";
  }
  if (s.dp_age) {
    nage = s.nage;
    if (nage <= 0 || nage > G) return "bad number of age groups";
    dp_age.assign(s.dp_age, s.dp_age + G);
    for (int a : dp_age)
      if (a < 0 || a >= nage) return "dp age group out of range";
  } else {
    nage = G;
    dp_age.resize(G);
    for (int g = 0; g < G; g++) dp_age[g] = g;
  }
  dirichlet.resize(s.ndirichlet);
  for (int k = 0; k < s.ndirichlet; k++) {
    dirichlet[k].param.assign(s.dirichlet[k].param, s.dirichlet[k].param + s.dirichlet[k].n);
    if (s.dirichlet[k].frac) dirichlet[k].frac.assign(s.dirichlet[k].frac, s.dirichlet[k].frac + s.dirichlet[k].n);
    else dirichlet[k].frac.assign(s.dirichlet[k].n, 1.0);
  }

  // Populations (src/data.cc:288-332,421-457).
  pop_init.assign(s.pop_init, s.pop_init + (size_t)A * C * G);
  area_pop.assign(A, 0.0);
  popsize = 0;
  std::vector<double> pop_age(G, 0.0);
  for (int c = 0; c < A; c++)
    for (int co = 0; co < C; co++)
      for (int g = 0; g < G; g++) {
        double p = pop_init[((size_t)c * C + co) * G + g];
        area_pop[c] += p;
        pop_age[g] += p;
      }
  for (int c = 0; c < A; c++) {
    if (area_pop[c] <= 0) return "area with zero population";
    popsize += area_pop[c];
  }
  Rdist.assign(G, 0.0);                                       // SplineInfo::democatpos_dist (src/model.cc:409-424)
  for (int g = 0; g < G; g++) Rdist[g] = pop_age[g] / popsize;

  // Age mixing: Data::age_mixing_matrix + agematrix_normalise
  // (src/data_matrix.cc:100-202), over the age groups' populations.
  std::vector<double> pop_a(nage, 0.0);
  for (int g = 0; g < G; g++) pop_a[dp_age[g]] += pop_age[g];
  N_age.assign((size_t)nage * nage, 0.0);
  if (s.N_raw) {
    std::vector<double> m((size_t)nage * nage);
    for (int j = 0; j < nage; j++)
      for (int i = 0; i < nage; i++) m[(size_t)j * nage + i] = s.N_raw[(size_t)j * nage + i] / (pop_a[j] / popsize);
    for (int j = 0; j < nage; j++)
      for (int i = 0; i < nage; i++)
        N_age[(size_t)j * nage + i] = 0.5 * (m[(size_t)j * nage + i] + m[(size_t)i * nage + j]);
  } else {
    if (nage != 1) return "an age mixing matrix is needed when there is more than one age group";
    N_age[0] = 1;
  }
  {
    double sum = 0, sum2 = 0;
    for (int a = 0; a < nage; a++)
      for (int aa = 0; aa < nage; aa++) { sum += pop_a[aa] * N_age[(size_t)aa * nage + a]; sum2 += pop_a[aa]; }
    sum /= sum2;
    for (auto& v : N_age) v /= sum;
  }
  // Between dps, the mixing of their age groups (State::get_NMI, src/state.cc:547-578).
  N.assign((size_t)G * G, 0.0);
  // With strains, a strain's infections mix with that strain's only.
  const int dpm = G / nstrain;
  for (int q = 0; q < G; q++)
    for (int d = 0; d < G; d++)
      N[(size_t)q * G + d] = q / dpm == d / dpm ? N_age[(size_t)dp_age[q] * nage + dp_age[d]] : 0.0;

  // Geographic mixing: Data::geo_normalise (src/data_matrix.cc:55-96).
  M_rowptr.assign(A + 1, 0);
  if (s.M_diag) {
    M_diag.assign(s.M_diag, s.M_diag + A);
    M_rowptr.assign(s.M_rowptr, s.M_rowptr + A + 1);
    int nnz = M_rowptr[A];
    M_to.assign(s.M_to, s.M_to + nnz);
    M_val.assign(s.M_val, s.M_val + nnz);
    for (int k = 0; k < nnz; k++)
      if (M_to[k] < 0 || M_to[k] >= A) return "geo mixing destination out of range";
  } else {
    if (A != 1) return "a geographic mixing matrix is needed when there is more than one area";
    M_diag.assign(1, 1.0);
  }
  M_factor.assign(A, 0.0);
  for (int c = 0; c < A; c++) {
    if (M_diag[c] == 0) return "the geographic mixing matrix cannot have zero on the diagonal";
    double sum = M_diag[c] * area_pop[c];
    for (int j = M_rowptr[c]; j < M_rowptr[c + 1]; j++) sum += M_val[j] * area_pop[M_to[j]];
    M_diag[c] /= sum;
    for (int j = M_rowptr[c]; j < M_rowptr[c + 1]; j++) M_val[j] /= sum;
  }
  for (int c = 0; c < A; c++) M_factor[c] = (1.0 / area_pop[c]) / M_diag[c];
  in_ptr.assign(A + 1, 0);
  for (int j = 0; j < M_rowptr[A]; j++) in_ptr[M_to[j] + 1]++;
  for (int c = 0; c < A; c++) in_ptr[c + 1] += in_ptr[c];
  in_src.assign(M_rowptr[A], 0);
  in_w.assign(M_rowptr[A], 0.0);
  {
    std::vector<int> fill(in_ptr.begin(), in_ptr.end() - 1);
    for (int c = 0; c < A; c++)
      for (int j = M_rowptr[c]; j < M_rowptr[c + 1]; j++) {
        int d = M_to[j];
        in_src[fill[d]] = c;
        in_w[fill[d]] = M_val[j];
        fill[d]++;
      }
  }

  // Covariates centred as Data::read_covars does (src/data.cc:557-565).
  covar.resize(s.ncovar);
  for (int k = 0; k < s.ncovar; k++) {
    covar[k].param = s.covar[k].param;
    covar[k].value.assign(s.covar[k].value, s.covar[k].value + (size_t)A * period);
    double av = 0;
    for (double v : covar[k].value) av += v;
    av /= (double)A * period;
    for (double& v : covar[k].value) v -= av;
    if (covar[k].param < 0 || covar[k].param >= nparam) return "covariate parameter out of range";
  }
  area_effect_param.clear();
  if (s.area_effect_param) {
    area_effect_param.assign(s.area_effect_param, s.area_effect_param + A);
    for (int th : area_effect_param)
      if (th < 0 || th >= nparam) return "area effect parameter out of range";
  }
  agemod.clear();
  for (int k = 0; k < s.nagemod; k++) {
    const beep_agemod& x = s.agemod[k];
    if (x.spline < 0 || x.spline >= s.nspline) return "age matrix modification spline out of range";
    AgeMod am{x.spline, std::vector<int>(x.ages, x.ages + x.nage)};
    for (int a : am.ages)
      if (a < 0 || a >= nage) return "age matrix modification age group out of range";
    agemod.push_back(am);
  }
  level_param.clear(); level_map.clear();
  if (s.nlevel > 0) {
    if (!s.level_param || !s.level_map) return "level effects need their parameters and map";
    level_param.assign(s.level_param, s.level_param + s.nlevel);
    level_map.assign(s.level_map, s.level_map + (size_t)period * A);
    for (int th : level_param)
      if (th < 0 || th >= nparam) return "level effect parameter out of range";
    for (int l : level_map)
      if (l < 0 || l >= s.nlevel) return "level effect map out of range";
  }

  // Infected states (Model::set_infected_uninfected, src/model.cc:110-166).
  {
    std::vector<char> fwd(C, 0), back(C, 0);
    std::vector<int> list;
    int c0 = trans[inf_tr].to;
    list.push_back(c0);
    fwd[c0] = 1;
    for (size_t i = 0; i < list.size(); i++) {
      int c = list[i];
      for (int tr : comp[c].exits) {
        if (trans[tr].infection) continue;
        int cc = trans[tr].to;
        if (!fwd[cc] && cc != trans[inf_tr].from) { list.push_back(cc); fwd[cc] = 1; }
      }
    }
    for (int co = 0; co < C; co++) {
      if (comp[co].inf_param < 0) continue;
      std::vector<int> l2(1, co);
      back[co] = 1;
      for (size_t i = 0; i < l2.size(); i++)
        for (int tr = 0; tr < T; tr++)
          if (!trans[tr].infection && trans[tr].to == l2[i] && !back[trans[tr].from]) {
            back[trans[tr].from] = 1;
            l2.push_back(trans[tr].from);
          }
    }
    inf_state.clear();
    for (int co = 0; co < C; co++)
      if (fwd[co] && back[co]) inf_state.push_back(co);
  }

  // Observations: ObservationModel::initialise_obs_change (src/obsmodel.cc:174-217)
  // flattened per observation, and Data::set_obs_sd (src/data.cc:1589-1623).
  obs.resize(s.nobs);
  obs_splines.clear();
  ndatatable = 1;
  std::vector<double> graph_max;
  for (int i = 0; i < s.nobs; i++) {
    const beep_obs& o = s.obs[i];
    if (o.graph >= (int)graph_max.size()) graph_max.resize(o.graph + 1, 0.0);
    if (o.y != BEEP_OBS_UNKNOWN && o.y != BEEP_OBS_THRESH && o.y > graph_max[o.graph]) graph_max[o.graph] = o.y;
  }
  for (int i = 0; i < s.nobs; i++) {
    const beep_obs& o = s.obs[i];
    Obs& ob = obs[i];
    ob.y = o.y; ob.shape = o.shape; ob.factor = o.factor; ob.thresh = o.thresh;
    ob.type = o.type; ob.datatable = o.datatable; ob.sett_i = o.sett_i; ob.sett_f = o.sett_f;
    if (o.type < 0 || o.type > 3) return "unknown observation model";
    if (o.sett_i < 0 || o.sett_f > S || o.sett_i >= o.sett_f) return "observation window out of range";
    if (o.datatable < 0) return "negative data table index";
    if (o.datatable + 1 > ndatatable) ndatatable = o.datatable + 1;
    ob.sd = o.sd;
    ob.ospl = -1;
    if (o.factor_spline != 0) {
      const int sp = o.factor_spline - 1;
      if (sp < 0 || sp >= s.nspline) return "observation spline out of range";
      if (o.nco > 0) return "an observation spline on a population observation";   // emsgEC("ObsModel",3)
      size_t k = 0;
      while (k < obs_splines.size() && obs_splines[k] != sp) k++;
      if (k == obs_splines.size()) obs_splines.push_back(sp);
      ob.ospl = (int)k;
    }
    if (o.type == BEEP_OBS_NORMAL_PERCENT && !(o.sd > 0))
      ob.sd = (o.y * (1 - EPSILON_FACTOR) + graph_max[o.graph] * EPSILON_FACTOR) * 0.01 * o.percent;
    if ((o.type == BEEP_OBS_NORMAL || o.type == BEEP_OBS_NORMAL_PERCENT) && !(ob.sd > 0) &&
        o.y != BEEP_OBS_UNKNOWN)
      return "normal observation with non-positive sd";
    for (int st = o.sett_i; st < o.sett_f; st++)
      for (int a = 0; a < o.narea; a++) {
        int c = o.area[a];
        if (c < 0 || c >= A) return "observation area out of range";
        for (int d = 0; d < o.ndp; d++) {
          int dp = o.dp[d];
          if (dp < 0 || dp >= G) return "observation dp out of range";
          uint32_t cell = (uint32_t)(c * G + dp);
          for (int k = 0; k < o.ntr; k++) {
            int tr = o.tr[k];
            if (tr < 0 || tr >= T) return "observation transition out of range";
            ob.elem.push_back((uint32_t)(((size_t)st * T + tr) * AG + cell));
          }
          for (int k = 0; k < o.nco; k++) {
            int co = o.co[k];
            if (co < 0 || co >= C) return "observation compartment out of range";
            ob.elem.push_back(0x80000000u | (uint32_t)(((size_t)st * C + co) * AG + cell));
          }
        }
      }
  }
  // ObservationModel::split_observations (src/obsmodel.cc:326-348).
  section_ti.assign(s.section_ti, s.section_ti + s.nsection);
  section_tf.assign(s.section_tf, s.section_tf + s.nsection);
  section_obs.assign(s.nsection, std::vector<int>());
  for (int k = 0; k < s.nsection; k++) {
    int ti = section_ti[k], tf = section_tf[k];
    if (ti < 0 || tf > S || ti >= tf) return "bad section";
    for (int i = 0; i < s.nobs; i++)
      if (obs[i].sett_i < tf && obs[i].sett_f >= ti && obs[i].sett_i >= ti && obs[i].sett_f <= tf)
        section_obs[k].push_back(i);
  }
  // The area factor depends on the parameters only through area effects,
  // covariates and level effects; without them it is one table for all.
  af_param = !area_effect_param.empty() || !covar.empty() || !level_param.empty();
  af_const.clear();
  if (!af_param) af_const = area_factor(std::vector<double>(nparam, 0.0));
  const std::string eg = build_groups(s);
  if (!eg.empty()) return eg;
  return build_modifications(s);
}

// The groups of R(t) and of the external force (Model::update_R_ref and
// update_efoi_ref, src/model.cc:384-480): each R group's dp distribution over
// its areas' initial populations (SplineInfo::democatpos_dist, the ratio's
// weights), each external-force group's spline and age weights.  Without
// groups in the spec: one of each, R_spline and efoi_spline everywhere.
std::string HostModel::build_groups(const beep_spec& s) {
  const int nspl = (int)spline.size();
  if (s.nrgroup < 0 || (s.nrgroup > 0 && (!s.rgroup_spline || !s.rgroup_area))) return "bad R-spline groups";
  if (s.negroup < 0 || (s.negroup > 0 && (!s.egroup_spline || !s.egroup_cell || !s.egroup_agedist)))
    return "bad external-force groups";
  if (s.nrgroup > 1 && nstrain > 1) return "R splines per area are not supported with strains";
  nR = s.nrgroup > 0 ? s.nrgroup : 1;
  if (s.nrgroup > 0) {
    rgrp_spline.assign(s.rgroup_spline, s.rgroup_spline + nR);
    rgrp_area.assign(s.rgroup_area, s.rgroup_area + A);
  } else {
    rgrp_spline.assign(1, R_spline);
    rgrp_area.assign(A, 0);
  }
  for (int sp : rgrp_spline)
    if (sp < 0 || sp >= nspl) return "R-spline group: spline out of range";
  std::vector<int> nareas(nR, 0);
  for (int r : rgrp_area) {
    if (r < 0 || r >= nR) return "R-spline group of an area out of range";
    nareas[r]++;
  }
  for (int n : nareas)
    if (n == 0) return "an R-spline group has no area";
  if (nR == 1) {
    rgrp_dist.assign(1, Rdist);
  } else {
    rgrp_dist.assign(nR, std::vector<double>(G, 0.0));
    std::vector<double> tot(nR, 0.0);
    for (int c = 0; c < A; c++)
      for (int co = 0; co < C; co++)
        for (int g = 0; g < G; g++) {
          const double p = pop_init[((size_t)c * C + co) * G + g];
          rgrp_dist[rgrp_area[c]][g] += p;
          tot[rgrp_area[c]] += p;
        }
    for (int r = 0; r < nR; r++) {
      if (!(tot[r] > 0)) return "an R-spline group's areas have no population";
      for (double& v : rgrp_dist[r]) v /= tot[r];
    }
  }
  nE = s.negroup > 0 ? s.negroup : 1;
  if (s.negroup > 0) {
    egrp_spline.assign(s.egroup_spline, s.egroup_spline + nE);
    egrp_cell.assign(s.egroup_cell, s.egroup_cell + (size_t)nstrain * A);
    egrp_agedist.assign(s.egroup_agedist, s.egroup_agedist + (size_t)nE * G);
    for (int sp : egrp_spline)
      if (sp < 0 || sp >= nspl) return "external-force group: spline out of range";
    for (int e : egrp_cell)
      if (e < 0 || e >= nE) return "external-force group of a cell out of range";
    for (double v : egrp_agedist)
      if (!(v >= 0) || std::isinf(v)) return "external-force age weights must be finite and non-negative";
  } else {
    egrp_spline.assign(1, efoi_spline);
    egrp_cell.assign((size_t)nstrain * A, 0);
    egrp_agedist = efoi_agedist;
  }
  return "";
}

// The tables of Model::model_modification's loop (src/model.cc:1495-1575):
// each modification multiplies its cells' factors, or sets a spline's steps;
// a spline step cannot be both set and multiplied (the reference's error).
std::string HostModel::build_modifications(const beep_spec& s) {
  mod_beta.clear();
  mod_efoi.clear();
  mod_trans.clear();
  spline_mult.clear();
  spline_set.clear();
  if (s.nmod < 0 || (s.nmod > 0 && !s.mod)) return "bad modification list";
  const int dpm = G / nstrain;
  for (int k = 0; k < s.nmod; k++) {
    const beep_mod& md = s.mod[k];
    if (md.start < 0 || md.end > S || md.start > md.end) return "modification: step range out of range";
    if (!(md.factor >= 0) || std::isinf(md.factor)) return "modification: the factor must be non-negative";
    std::vector<int> areas, dps;
    if (md.narea > 0) {
      if (!md.area) return "modification: no area list";
      for (int i = 0; i < md.narea; i++) {
        if (md.area[i] < 0 || md.area[i] >= A) return "modification: area out of range";
        areas.push_back(md.area[i]);
      }
    } else {
      for (int c = 0; c < A; c++) areas.push_back(c);
    }
    if (md.ndp > 0) {
      if (!md.dp) return "modification: no dp list";
      for (int i = 0; i < md.ndp; i++) {
        if (md.dp[i] < 0 || md.dp[i] >= G) return "modification: dp out of range";
        dps.push_back(md.dp[i]);
      }
    } else {
      for (int g = 0; g < G; g++) dps.push_back(g);
    }
    switch (md.type) {
      case BEEP_MOD_TRANS_RATE: {
        if (md.ntr <= 0 || !md.tr) return "trans_rate_fac: no transitions";
        if (mod_trans.empty()) mod_trans.assign((size_t)S * T * AG, 1.0);
        for (int i = 0; i < md.ntr; i++) {
          const int tr = md.tr[i];
          if (tr < 0 || tr >= T) return "trans_rate_fac: transition out of range";
          for (int st = md.start; st < md.end; st++)
            for (int c : areas)
              for (int g : dps) mod_trans[((size_t)st * T + tr) * AG + (size_t)c * G + g] *= md.factor;
        }
        break;
      }
      case BEEP_MOD_BETA:
      case BEEP_MOD_EFOI: {
        if (md.strain < 0 || md.strain >= nstrain) return "beta_fac / efoi_fac: strain out of range";
        std::vector<double>& tab = md.type == BEEP_MOD_BETA ? mod_beta : mod_efoi;
        if (tab.empty()) tab.assign((size_t)S * AG, 1.0);
        // The strain's dps: its infections (src/state.cc:409-430).
        for (int st = md.start; st < md.end; st++)
          for (int c : areas)
            for (int g = md.strain * dpm; g < (md.strain + 1) * dpm; g++) tab[(size_t)st * AG + (size_t)c * G + g] *= md.factor;
        break;
      }
      case BEEP_MOD_SPLINE_FAC:
      case BEEP_MOD_SPLINE_SET: {
        if (md.spline < 0 || md.spline >= (int)spline.size()) return "spline_fac / spline_set: spline out of range";
        if (spline_mult.empty()) {
          spline_mult.assign(spline.size(), std::vector<double>(S, 1.0));
          spline_set.assign(spline.size(), std::vector<double>(S, std::nan("")));
        }
        for (int st = md.start; st < md.end; st++) {
          double& mult = spline_mult[md.spline][st];
          double& set = spline_set[md.spline][st];
          if (md.type == BEEP_MOD_SPLINE_FAC) {
            if (!std::isnan(set)) return "In 'modification' a spline cannot both be directly set and have a factor change";
            mult *= md.factor;
          } else {
            if (mult != 1) return "In 'modification' a spline cannot both be directly set and have a factor change";
            set = md.factor;
          }
        }
        break;
      }
      default:
        return "modification: unknown type";
    }
  }
  return "";
}

std::vector<double> HostModel::dirichlet_correct(const double* paramv) const {
  std::vector<double> p(paramv, paramv + nparam);
  for (const auto& d : dirichlet) {                          // src/model.cc:2685-2695
    double sum = 0;
    for (int th : d.param) sum += paramv[th];
    for (size_t k = 0; k < d.param.size(); k++) p[d.param[k]] = paramv[d.param[k]] / (d.frac[k] * sum);
  }
  return p;
}

// Model::create_disc_spline (src/model.cc:1020-1053), SIM-mode hold past the last knot.
std::vector<double> HostModel::disc_spline(int sp, const std::vector<double>& pd) const {
  std::vector<double> out(S, 0.0);
  const Spline& spl = spline[sp];
  double factor = spl.factor_param >= 0 ? pd[spl.factor_param] : 1.0;
  size_t p = 0;
  for (int s = 0; s < S; s++) {
    double t = (s + 0.5) * period / S;
    while (p + 1 < spl.t.size() && t > spl.t[p + 1]) p++;
    if (p + 1 == spl.t.size()) {
      out[s] = s > 0 ? out[s - 1] : 0.0;
    } else {
      double fac = (t - spl.t[p]) / (spl.t[p + 1] - spl.t[p]);
      out[s] = factor * (pd[spl.param[p]] * spl.multfac[p] * (1 - fac) + pd[spl.param[p + 1]] * spl.multfac[p + 1] * fac);
    }
  }
  // Prediction-mode modifications (src/model.cc:1044-1049).
  if (!spline_mult.empty())
    for (int s = 0; s < S; s++) {
      out[s] *= spline_mult[sp][s];
      if (!std::isnan(spline_set[sp][s])) out[s] = spline_set[sp][s];
    }
  return out;
}

int HostModel::derive(const double* paramv, Tables& tb, std::string& err) const {
  std::vector<double> pd = dirichlet_correct(paramv);

  // Model::create_transrate (src/model.cc:1116-1140).
  tb.rate.assign((size_t)T * G, 0.0);
  for (int t = 0; t < T; t++) {
    if (trans[t].infection) continue;
    const Compartment& c = comp[trans[t].from];
    for (int g = 0; g < G; g++) {
      double mean = pd[c.mean_param[g]] / c.shape;
      if (mean == 0) { err = "zero mean residence time"; return BEEP_EINVAL; }
      double r = 1.0 / mean;
      if (c.exits.size() > 1) r *= pd[trans[t].prob_param[g]];
      tb.rate[(size_t)t * G + g] = r;
    }
  }
  // Model::create_susceptibility (src/model.cc:1144-1178): the product over categories.
  tb.sus.assign(G, 1.0);
  if (!sus_param.empty())
    for (int g = 0; g < G; g++) tb.sus[g] = sus_param[g] >= 0 ? pd[sus_param[g]] : 1.0;
  for (const auto& f : sus_extra)
    for (int g = 0; g < G; g++)
      if (f[g] >= 0) tb.sus[g] *= pd[f[g]];
  tb.inf.assign(C, 0.0);
  for (int c = 0; c < C; c++) tb.inf[c] = comp[c].inf_param >= 0 ? pd[comp[c].inf_param] : 0.0;
  // Model::create_Ntime (src/model.cc:1254-1301): each modification
  // multiplies its age groups' rows and columns by its spline's value.
  // The age matrix is modified, then spread over the dps.
  tb.Ntime = N;
  if (!agemod.empty()) {
    std::vector<std::vector<double>> fac;
    for (const AgeMod& am : agemod) fac.push_back(disc_spline(am.spline, pd));
    tb.Ntime.assign((size_t)S * G * G, 0.0);
    std::vector<double> Na(N_age.size());
    for (int s = 0; s < S; s++) {
      Na = N_age;
      for (size_t k = 0; k < agemod.size(); k++) {
        const double f = fac[k][s];
        for (int a : agemod[k].ages)
          for (int i = 0; i < nage; i++) {
            Na[(size_t)a * nage + i] *= f;
            Na[(size_t)i * nage + a] *= f;
          }
      }
      double* Ns = tb.Ntime.data() + (size_t)s * G * G;
      const int dpm = G / nstrain;
      for (int q = 0; q < G; q++)
        for (int d = 0; d < G; d++)
          Ns[(size_t)q * G + d] = q / dpm == d / dpm ? Na[(size_t)dp_age[q] * nage + dp_age[d]] : 0.0;
    }
  }

  if (af_param) tb.af = area_factor(pd);
  else tb.af.clear();

  // Model::calculate_beta_from_R (src/model.cc:1318-1362) with calculate_Vinv,
  // calculate_F, calculate_NGM and the dominant eigenvalue (2485-2621).
  if (nstrain > 1) {
    int rc = strain_beta(pd, tb, err);
    if (rc) return rc;
  } else {
    const int nis = (int)inf_state.size();
    if (nis == 0) { err = "no infected states"; return BEEP_EINVAL; }
    std::vector<int> ref(C, -1);
    for (int i = 0; i < nis; i++) ref[inf_state[i]] = i;
    const int NV = nis * G;
    std::vector<double> V((size_t)NV * NV, 0.0), Vinv;
    for (int t = 0; t < T; t++) {
      int fi = ref[trans[t].from], ti = ref[trans[t].to];
      if (fi < 0) continue;
      for (int g = 0; g < G; g++) {
        double r = tb.rate[(size_t)t * G + g];
        int j = fi * G + g;
        V[(size_t)j * NV + j] += r;
        if (ti >= 0) V[(size_t)(ti * G + g) * NV + j] -= r;
      }
    }
    if (!invert(V, NV, Vinv)) { err = "singular transition matrix V"; return BEEP_EINVAL; }
    // calculate_R_beta_ratio_using_NGM: dominant eigenvalue of F V^-1 for one
    // age contact matrix Nm (src/model.cc:2485-2576).
    auto ngm_ratio = [&](const double* Nm, const std::vector<double>& dist) {
      std::vector<double> F((size_t)G * NV, 0.0);
      for (int si = 0; si < nis; si++) {
        double inf = tb.inf[inf_state[si]];
        if (inf == 0) continue;
        for (int dp = 0; dp < G; dp++) {
          int i = si * G + dp;
          for (int q = 0; q < G; q++) F[(size_t)q * NV + i] += dist[q] * tb.sus[q] * inf * Nm[(size_t)q * G + dp];
        }
      }
      std::vector<double> NGM((size_t)G * G, 0.0);
      for (int q = 0; q < G; q++)
        for (int i = 0; i < G; i++) {
          double sum = 0;
          for (int j = 0; j < NV; j++) sum += F[(size_t)q * NV + j] * Vinv[(size_t)j * NV + i];
          NGM[(size_t)q * G + i] = sum;
        }
      return dominant_eigenvalue(NGM, G);
    };
    // The ratio is recomputed only where the matrix differs from the previous
    // step's by more than TINY = 1e-8 in some entry (src/model.cc:1337-1339,1303-1314).
    const bool tv = !agemod.empty();
    const double Rfac = Rfactor_param >= 0 ? pd[Rfactor_param] : 1.0;
    tb.beta.assign((size_t)S * nR, 0.0);
    // Each R group's ratio over its own areas' dp distribution.
    for (int r = 0; r < nR; r++) {
      std::vector<double> ratio(tv ? S : 1);
      for (int s = 0; s < (int)ratio.size(); s++) {
        const double* Ns = tb.Ntime.data() + (size_t)s * G * G;
        bool same = s > 0;
        for (int k = 0; same && k < G * G; k++) {
          const double v1 = Ns[k - G * G], v2 = Ns[k];
          if (v1 < v2 - 1e-8 || v1 > v2 + 1e-8) same = false;
        }
        ratio[s] = same ? ratio[s - 1] : ngm_ratio(Ns, rgrp_dist[r]);
        if (!(ratio[s] > 0) || std::isnan(ratio[s])) {
          err = "next-generation matrix has no positive dominant eigenvalue";
          return BEEP_EINVAL;
        }
      }
      const std::vector<double> Rt = disc_spline(rgrp_spline[r], pd);
      for (int s = 0; s < S; s++) tb.beta[(size_t)s * nR + r] = Rfac * Rt[s] / ratio[tv ? s : 0];
    }
  }
  tb.efoi.assign((size_t)S * nE, 0.0);
  for (int e = 0; e < nE; e++) {
    if (egrp_spline[e] < 0) continue;
    const std::vector<double> v = disc_spline(egrp_spline[e], pd);
    for (int s = 0; s < S; s++) tb.efoi[(size_t)s * nE + e] = v[s] / popsize;
  }
  tb.geo.assign(S, 1.0);
  if (geo_spline >= 0) tb.geo = disc_spline(geo_spline, pd);
  // State::disc_spline of the observation splines (src/state.cc:76).
  tb.obs_spl.clear();
  for (int sp : obs_splines) {
    std::vector<double> v = disc_spline(sp, pd);
    tb.obs_spl.insert(tb.obs_spl.end(), v.begin(), v.end());
  }
  for (double v : tb.beta) if (std::isnan(v)) { err = "NaN transmission rate"; return BEEP_ENAN; }
  for (double v : tb.af) if (std::isnan(v)) { err = "NaN area factor"; return BEEP_ENAN; }
  return BEEP_OK;
}

// Model::create_areafactor (src/model.cc:1182-1250): area effects, covariates
// and level effects, normalised by the population-weighted time average.
std::vector<double> HostModel::area_factor(const std::vector<double>& pd) const {
  std::vector<double> af((size_t)period * A, 1.0);
  if (!area_effect_param.empty())
    for (int t = 0; t < period; t++)
      for (int c = 0; c < A; c++) af[(size_t)t * A + c] = pd[area_effect_param[c]];
  for (const auto& cv : covar) {
    double b = pd[cv.param];
    for (int t = 0; t < period; t++)
      for (int c = 0; c < A; c++) af[(size_t)t * A + c] *= std::exp(b * cv.value[(size_t)c * period + t]);
  }
  if (!level_param.empty())
    for (size_t i = 0; i < af.size(); i++) af[i] *= pd[level_param[level_map[i]]];
  double area_av = 0;
  for (int c = 0; c < A; c++) {
    double av = 0;
    for (int t = 0; t < period; t++) av += af[(size_t)t * A + c];
    av /= period;
    area_av += av * area_pop[c];
  }
  area_av /= popsize;
  for (auto& v : af) v /= area_av;
  return af;
}

std::vector<double> HostModel::efoi_cells(const Tables& tb) const {
  const int dpm = G / nstrain;
  std::vector<double> out((size_t)S * A * G);
  for (int s = 0; s < S; s++)
    for (int c = 0; c < A; c++)
      for (int g = 0; g < G; g++) {
        const int e = egrp_cell[(size_t)(g / dpm) * A + c];
        out[((size_t)s * A + c) * G + g] = tb.efoi[(size_t)s * nE + e] * egrp_agedist[(size_t)e * G + g];
      }
  return out;
}

std::vector<double> HostModel::beta_af(const Tables& tb) const {
  const std::vector<double>& af = af_param ? tb.af : af_const;
  std::vector<double> out((size_t)S * A, 0.0);
  for (int s = 0; s < S; s++)
    for (int c = 0; c < A; c++)
      out[(size_t)s * A + c] = tb.beta[(size_t)s * nR + rgrp_area[c]] * af[(size_t)(s / spd) * A + c];
  return out;
}

// Strains (Model::calculate_beta_from_R, src/model.cc:1318-1348, with
// calculate_Vinv and calculate_F per strain): each strain's NGM ratio over its
// own dps, with the population fractions pooled over strains (democatpos_dist,
// src/data.cc:154) and that strain's susceptibility; beta_st = Rfactor_st *
// R(t) / ratio_st.  The device takes one beta per step for all groups, so
// the strain factor Rfactor_st / ratio_st goes into the strain's block of the
// (block-diagonal) contact matrix, and beta holds R(t).
int HostModel::strain_beta(const std::vector<double>& pd, Tables& tb, std::string& err) const {
  const int nis = (int)inf_state.size();
  if (nis == 0) { err = "no infected states"; return BEEP_EINVAL; }
  const int dpm = G / nstrain, NV = nis * dpm;
  std::vector<int> ref(C, -1);
  for (int i = 0; i < nis; i++) ref[inf_state[i]] = i;
  std::vector<double> Rd(dpm, 0.0);
  for (int g = 0; g < G; g++) Rd[g % dpm] += Rdist[g];
  const bool tv = !agemod.empty();
  const int nsteps = tv ? S : 1;
  std::vector<double> fold((size_t)nsteps * nstrain);
  for (int st = 0; st < nstrain; st++) {
    std::vector<double> V((size_t)NV * NV, 0.0), Vinv;
    for (int t = 0; t < T; t++) {
      int fi = ref[trans[t].from], ti = ref[trans[t].to];
      if (fi < 0) continue;
      for (int q = 0; q < dpm; q++) {
        double r = tb.rate[(size_t)t * G + st * dpm + q];
        int j = fi * dpm + q;
        V[(size_t)j * NV + j] += r;
        if (ti >= 0) V[(size_t)(ti * dpm + q) * NV + j] -= r;
      }
    }
    if (!invert(V, NV, Vinv)) { err = "singular transition matrix V"; return BEEP_EINVAL; }
    const double Rfac = strain_Rfactor[st] >= 0 ? pd[strain_Rfactor[st]] : 1.0;
    double ratio = 0;
    for (int s = 0; s < nsteps; s++) {
      const double* Ns = tb.Ntime.data() + (size_t)s * G * G;
      bool same = s > 0;
      for (int k = 0; same && k < G * G; k++) {
        const double v1 = Ns[k - G * G], v2 = Ns[k];
        if (v1 < v2 - 1e-8 || v1 > v2 + 1e-8) same = false;
      }
      if (!same) {
        std::vector<double> F((size_t)dpm * NV, 0.0);
        for (int si = 0; si < nis; si++) {
          double inf = tb.inf[inf_state[si]];
          if (inf == 0) continue;
          for (int d = 0; d < dpm; d++)
            for (int q = 0; q < dpm; q++)
              F[(size_t)q * NV + si * dpm + d] +=
                  Rd[q] * tb.sus[st * dpm + q] * inf * Ns[(size_t)(st * dpm + q) * G + st * dpm + d];
        }
        std::vector<double> NGM((size_t)dpm * dpm, 0.0);
        for (int q = 0; q < dpm; q++)
          for (int i = 0; i < dpm; i++) {
            double sum = 0;
            for (int j = 0; j < NV; j++) sum += F[(size_t)q * NV + j] * Vinv[(size_t)j * NV + i];
            NGM[(size_t)q * dpm + i] = sum;
          }
        ratio = dominant_eigenvalue(NGM, dpm);
        if (!(ratio > 0) || std::isnan(ratio)) {
          err = "next-generation matrix has no positive dominant eigenvalue";
          return BEEP_EINVAL;
        }
      }
      fold[(size_t)s * nstrain + st] = Rfac / ratio;
    }
  }
  for (int s = 0; s < nsteps; s++) {
    double* Ns = tb.Ntime.data() + (size_t)s * G * G;
    for (int q = 0; q < G; q++)
      for (int d = 0; d < G; d++) Ns[(size_t)q * G + d] *= fold[(size_t)s * nstrain + q / dpm];
  }
  tb.beta = disc_spline(R_spline, pd);
  return BEEP_OK;
}

double HostModel::prior_logpdf(const double* pv) const {   // Model::prior (src/model.cc:1590-1659)
  double Pr = 0;
  for (int th = 0; th < nparam; th++) {
    const beep_prior& p = prior[th];
    switch (p.type) {
      case BEEP_PRIOR_UNIFORM: Pr += std::log(1.0 / (p.b - p.a)); break;
      case BEEP_PRIOR_EXP: Pr += std::log(p.a) - p.a * pv[th]; break;
      case BEEP_PRIOR_NORMAL:
        Pr += -0.5 * std::log(2 * M_PI * p.b * p.b) - (pv[th] - p.a) * (pv[th] - p.a) / (2 * p.b * p.b);
        break;
      case BEEP_PRIOR_DIRFLAT: Pr += -pv[th]; break;
      case BEEP_PRIOR_GAMMA:                                 // gamma_probability, src/utils.cc:115-120
        Pr += (p.a - 1) * std::log(pv[th]) - p.b * pv[th] + p.a * std::log(p.b) - std::lgamma(p.a);
        break;
      default: break;
    }
  }
  // Model::spline_prior (src/model.cc:1662-1704), on the Dirichlet-corrected values.
  std::vector<double> pd;
  for (const Spline& spl : spline) {
    if (spl.smooth_type == BEEP_SMOOTH_NONE) continue;
    if (pd.empty()) pd = dirichlet_correct(pv);
    for (size_t i = 0; i + 1 < spl.param.size(); i++) {
      int p1 = spl.param[i], p2 = spl.param[i + 1];
      if (p1 == p2 || spl.t[i] == spl.t[i + 1]) continue;
      if (prior[p1].type == BEEP_PRIOR_FIXED && prior[p2].type == BEEP_PRIOR_FIXED) continue;
      const double v = spl.smooth[i];
      if (spl.smooth_type == BEEP_SMOOTH_LOG) {
        const double mu = std::log(pd[p1] * spl.multfac[i]), logx = std::log(pd[p2] * spl.multfac[i + 1]);
        Pr += -(logx - mu) * (logx - mu) / (2 * v * v) - logx;
      } else {
        const double d = pd[p1] - pd[p2];
        Pr += -d * d / (2 * v * v);
      }
    }
  }
  return Pr;
}

bool HostModel::inbounds(const double* pv) const {         // Model::inbounds (src/model.cc:2625-2669)
  for (int th = 0; th < nparam; th++) {
    const beep_prior& p = prior[th];
    double v = pv[th];
    switch (p.type) {
      case BEEP_PRIOR_FIXED: if (v != p.a) return false; break;
      case BEEP_PRIOR_UNIFORM: if (v < p.a || v > p.b) return false; break;
      case BEEP_PRIOR_EXP: case BEEP_PRIOR_DIRFLAT: case BEEP_PRIOR_GAMMA: if (v < 0) return false; break;
      default: break;
    }
  }
  return true;
}

}  // namespace beephost
