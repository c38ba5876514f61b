// TEST INFRASTRUCTURE ONLY.  This file is the FP64 CPU oracle for the BEEP
// hot path: a restatement of the reference's algorithm used by tests/, by the
// smoke check and by bench.py's cpu_baseline leg to CHECK the HIP product.
// Nothing in beep_amd/ may link, load or call it.
//
// It restates, over flat arrays in the reference's own index order, the
// functions of the simulation/likelihood path:
//   State::simulate(ti,tf)          src/state.cc:596-638
//   State::set_I_from_pop           src/state.cc:223-290
//   State::update_I_from_transnum   src/state.cc:178-219
//   State::set_Imap_sett            src/state.cc:164-174
//   State::get_NMI                  src/state.cc:547-578
//   State::set_transmean            src/state.cc:380-456
//   State::update_pop               src/state.cc:460-485
//   State::initialise_from_particle src/state.cc:294-322 (replay of a trajectory)
//   State::likelihood               src/state.cc:812-826
//   ObservationModel::get_obs_value_section / obs_prob
//                                   src/obsmodel.cc:74-109, 245-305
//   Mbp::mbp                        src/mbp.cc:252-330
//   Mbp::get_al                     src/mbp.cc:334-361
//   poisson_sample, binomial_sample src/utils.cc:148-153, 174-180
//   normal/poisson/negative_binomial_probability
//                                   src/utils.cc:107-127, 157-170
//
// Layouts (reference order, one strain, dp == age group):
//   pop[S][A][C][G], transnum/transmean[S][A][T][G], Imap/Idiag[S][A*G].
// The last simulated step draws transitions but does not update pop
// (src/state.cc:634), so pop has S slots.
//
// Samplers are libstdc++ std::poisson_distribution / binomial_distribution on
// std::default_random_engine (minstd_rand0), constructed per call exactly as
// src/utils.cc:148-153,174-180 do, so an oracle seeded like sran() produces
// the reference's own sample stream.
//
// Parity pins (each checked by a test under tests/): sampler streams and the
// log-pmfs against the reference's own utils.cc compiled into oracle/_ref;
// update_pop by the exact integer identity and conservation; the contraction
// by brute-force dense products and by "incremental Imap == Imap recomputed
// from pop" (the reference's own State::check invariant, state_check.cc:54-58);
// set_transmean and the dynamics by the forward-Euler SEIR recursion in
// deterministic mode and the closed form FOI = beta*I/N for one area;
// log-pmfs against scipy.stats; Mbp::mbp by the identity case (unchanged
// parameters give n_p == n_i) and by thinning/superposition laws.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

namespace {

const double LARGE = 1000000000.0;                 // src/consts.hh:174
const double UNKNOWN_OBS = 999999997.0;            // src/consts.hh:178 (UNKNOWN)
const double THRESH_OBS = 999999998.0;             // src/consts.hh:177 (THRESH)

std::default_random_engine g_gen;                  // src/utils.cc:34

int poisson_sample(double lam) {                   // src/utils.cc:148-153
  if (lam > LARGE) return -1;
  std::poisson_distribution<int> d(lam);
  return d(g_gen);
}

unsigned binomial_sample(double p, unsigned n) {   // src/utils.cc:174-180
  if (n == 0) return 0;
  std::binomial_distribution<int> d(n, p);
  return d(g_gen);
}

double normal_probability(double x, double mean, double var) {   // utils.cc:107-111
  return -0.5 * std::log(2 * M_PI * var) - (x - mean) * (x - mean) / (2 * var);
}

double poisson_probability(int i, double lam) {    // src/utils.cc:157-170
  if (lam == 0) return i == 0 ? 0.0 : -LARGE;
  if (i < 0) return -LARGE;
  return i * std::log(lam) - lam - std::lgamma(i + 1);
}

double negative_binomial_probability(unsigned val, double m, double shape) {  // utils.cc:124-127
  return std::lgamma(shape + val) - std::lgamma(val + 1) - std::lgamma(shape) +
         shape * std::log(shape / (shape + m)) + val * std::log(m / (shape + m));
}

}  // namespace

extern "C" {

// Topology after Data::geo_normalise (src/data_matrix.cc:55-96).
struct OrModel {
  int A, G, C, T, S;
  int spd;             // details.division_per_time
  int inf_tr;          // model.infection_trans
  int stochastic;      // details.stochastic
  double dt;           // period / ndivision
  const int* from;     // [T]
  const int* to;       // [T]
  const double* Mdiag; // [A]   genQ.M.diag
  const int* Mptr;     // [A+1] rows of genQ.M.to / genQ.M.val
  const int* Mto;      // [nnz]
  const double* Mval;  // [nnz]
  const double* Mfac;  // [A]   genQ.factor
  int nstrain;         // data.nstrain (1 = one strain); strains are the last category
};

// Quantities State::set_param derives from one parameter vector.
struct OrTables {
  const double* inf;     // [C]      paramval[comp[co].infectivity_param]
  const double* rate;    // [T][G]   transrate (unused for the infection transition)
  const double* sus;     // [G]      susceptibility
  const double* beta;    // [S][A]   beta[0][c][s]
  const double* af;      // [days][A] areafactor
  const double* efoi;    // [S]      disc_spline[efoi][s] / popsize
  const double* agedist; // [G]      efoi_agedist
  const double* geo;     // [S]      disc_spline[geo_spline_ref][s]
  const double* Ntime;   // [S][G][G]
  // Prediction-mode modifications (Model::modelmod, src/struct.hh:488-495),
  // null when absent: beta_mult and efoi_mult of each dp's strain, [S][A][G],
  // and transmean_mult, [S][A][T][G].
  const double* beta_mult;
  const double* efoi_mult;
  const double* transmean_mult;
  // External-force splines per strain and area (efoi_spl_ref[st][c],
  // src/state.cc:410-424): each cell's efoi times its age weight, [S][A][G];
  // null = efoi[s] * agedist[dp] everywhere.
  const double* efoi_cell;
};

struct OrState {
  double* pop;        // [S][A][C][G]
  double* transnum;   // [S][A][T][G]
  double* transmean;  // [S][A][T][G]
  double* Imap;       // [S][A*G]
  double* Idiag;      // [S][A*G]
};

void or_seed(unsigned seed) { g_gen.seed(seed); }

static inline size_t ix_pop(const OrModel* m, int s, int c, int co, int g) {
  return (((size_t)s * m->A + c) * m->C + co) * m->G + g;
}
static inline size_t ix_tr(const OrModel* m, int s, int c, int tr, int g) {
  return (((size_t)s * m->A + c) * m->T + tr) * m->G + g;
}

// State::set_I_from_pop (src/state.cc:223-290)
void or_set_I_from_pop(const OrModel* m, const OrTables* tb, OrState* st, int s) {
  const int A = m->A, G = m->G, C = m->C;
  double* Ima = st->Imap + (size_t)s * A * G;
  double* Idia = st->Idiag + (size_t)s * A * G;
  for (int v = 0; v < A * G; v++) { Ima[v] = 0; Idia[v] = 0; }
  std::vector<double> dinf(G);
  for (int c = 0; c < A; c++) {
    for (int a = 0; a < G; a++) dinf[a] = 0;
    for (int co = 0; co < C; co++) {
      double inf = tb->inf[co];
      if (inf != 0) {
        for (int dp = 0; dp < G; dp++) {
          double p = st->pop[ix_pop(m, s, c, co, dp)];
          if (p != 0) dinf[dp] += inf * p;
        }
      }
    }
    double diag = m->Mdiag[c];
    int v = c * G;
    for (int a = 0; a < G; a++) {
      double di = dinf[a];
      if (di != 0) {
        Idia[v + a] += di * diag;
        for (int j = m->Mptr[c]; j < m->Mptr[c + 1]; j++) Ima[m->Mto[j] * G + a] += di * m->Mval[j];
      }
    }
  }
}

// State::update_I_from_transnum (src/state.cc:178-219), applied to the arrays
// Ima/Idia with the per-step transition numbers dtn[A][T][G].
static void update_I_from_transnum(const OrModel* m, const OrTables* tb, double* Ima, double* Idia,
                                   const double* dtn) {
  const int A = m->A, G = m->G, T = m->T;
  std::vector<double> dinf(G);
  for (int c = 0; c < A; c++) {
    for (int a = 0; a < G; a++) dinf[a] = 0;
    for (int tr = 0; tr < T; tr++) {
      double di = tb->inf[m->to[tr]] - tb->inf[m->from[tr]];   // Model::get_infectivity_dif
      if (di != 0) {
        for (int dp = 0; dp < G; dp++) {
          double num = dtn[((size_t)c * T + tr) * G + dp];
          if (num != 0) dinf[dp] += di * num;
        }
      }
    }
    double diag = m->Mdiag[c];
    int v = c * G;
    for (int a = 0; a < G; a++) {
      double di = dinf[a];
      if (di != 0) {
        Idia[v + a] += di * diag;
        for (int j = m->Mptr[c]; j < m->Mptr[c + 1]; j++) Ima[m->Mto[j] * G + a] += di * m->Mval[j];
      }
    }
  }
}

// State::set_Imap_sett (src/state.cc:164-174)
void or_set_Imap_sett(const OrModel* m, const OrTables* tb, OrState* st, int s) {
  if (s == 0) { or_set_I_from_pop(m, tb, st, 0); return; }
  const size_t AG = (size_t)m->A * m->G;
  std::memcpy(st->Imap + s * AG, st->Imap + (s - 1) * AG, AG * sizeof(double));
  std::memcpy(st->Idiag + s * AG, st->Idiag + (s - 1) * AG, AG * sizeof(double));
  update_I_from_transnum(m, tb, st->Imap + s * AG, st->Idiag + s * AG,
                         st->transnum + (size_t)(s - 1) * m->A * m->T * m->G);
}

// State::get_NMI (src/state.cc:547-578)
static void get_NMI(const OrModel* m, const OrTables* tb, const OrState* st, int s, int c, double* NMI) {
  const int G = m->G;
  double d = tb->geo[s], omd = 1 - d, fac = m->Mfac[c];
  const double* Idg = st->Idiag + (size_t)s * m->A * G;
  const double* Imp = st->Imap + (size_t)s * m->A * G;
  if (G == 1) {
    double val = Idg[c] * (d + fac * omd) + Imp[c] * d;
    if (val < 0) val = 0;
    NMI[0] = val;
    return;
  }
  std::vector<double> I(G);
  int v = c * G;
  for (int a = 0; a < G; a++) I[a] = (Idg[v + a] + Imp[v + a]) * d + fac * Idg[v + a] * omd;
  const double* N = tb->Ntime + (size_t)s * G * G;
  for (int a = 0; a < G; a++) {
    double sum = 0;
    for (int aa = 0; aa < G; aa++) sum += N[a * G + aa] * I[aa];
    if (sum < 0) sum = 0;
    NMI[a] = sum;
  }
}

// State::set_transmean (src/state.cc:380-456), one strain.  Returns 0, or -4
// when a mean is NaN (the reference's emsgEC("State",4)).
int or_set_transmean(const OrModel* m, const OrTables* tb, OrState* st, int s, int c) {
  const int G = m->G, T = m->T;
  double dt = m->dt;
  std::vector<double> NMI(G);
  get_NMI(m, tb, st, s, c, NMI.data());
  int tr = m->inf_tr;
  const double beta_sc = tb->beta[(size_t)s * m->A + c];
  const double af_sc = tb->af[(size_t)(s / m->spd) * m->A + c];
  double be = beta_sc * af_sc;
  double et = tb->efoi[s];
  // Strains (src/state.cc:395-404): the susceptibles of a dp are pooled into
  // strain 0 (in the state itself) and every strain's infection draws on the
  // pool.  (The tables carry each strain's Rfactor / ratio in its block of
  // Ntime, so be is R(t) * areafactor here.)
  const int ns = m->nstrain > 1 ? m->nstrain : 1, dpm = G / ns;
  if (ns > 1)
    for (int dp = 0; dp < dpm; dp++) {
      double sum = 0;
      for (int k = 1; k < ns; k++) {
        double& v = st->pop[ix_pop(m, s, c, m->from[tr], k * dpm + dp)];
        sum += v;
        v = 0;
      }
      st->pop[ix_pop(m, s, c, m->from[tr], dp)] += sum;
    }
  for (int dp = 0; dp < G; dp++) {
    double popu = st->pop[ix_pop(m, s, c, m->from[tr], dp % dpm)];
    double& out = st->transmean[ix_tr(m, s, c, tr, dp)];
    if (popu <= 0) out = 0;
    else {
      // beta[st][c][sett] carries beta_mult (src/model.cc:1349-1356) and et
      // efoi_mult (src/state.cc:415).
      double b = be, e = et;
      if (tb->beta_mult) b = beta_sc * tb->beta_mult[((size_t)s * m->A + c) * G + dp] * af_sc;
      if (tb->efoi_mult) e = et * tb->efoi_mult[((size_t)s * m->A + c) * G + dp];
      double eterm = e * tb->agedist[dp];
      if (tb->efoi_cell) {
        eterm = tb->efoi_cell[((size_t)s * m->A + c) * G + dp];
        if (tb->efoi_mult) eterm *= tb->efoi_mult[((size_t)s * m->A + c) * G + dp];
      }
      out = dt * popu * tb->sus[dp] * (b * NMI[dp] + eterm);
      if (std::isnan(out)) return -4;
    }
  }
  for (int t = 0; t < T; t++) {
    if (t == m->inf_tr) continue;
    int from = m->from[t];
    for (int dp = 0; dp < G; dp++) {
      double popu = st->pop[ix_pop(m, s, c, from, dp)];
      st->transmean[ix_tr(m, s, c, t, dp)] = popu <= 0 ? 0 : dt * popu * tb->rate[t * G + dp];
    }
  }
  // transmean_mult (src/state.cc:446-453).
  if (tb->transmean_mult)
    for (int t = 0; t < T; t++)
      for (int dp = 0; dp < G; dp++) st->transmean[ix_tr(m, s, c, t, dp)] *= tb->transmean_mult[ix_tr(m, s, c, t, dp)];
  return 0;
}

// State::update_pop (src/state.cc:460-485): pop[s+1] = pop[s] - out + in.
void or_update_pop(const OrModel* m, OrState* st, int s) {
  const int A = m->A, C = m->C, G = m->G, T = m->T;
  std::memcpy(st->pop + ix_pop(m, s + 1, 0, 0, 0), st->pop + ix_pop(m, s, 0, 0, 0),
              sizeof(double) * (size_t)A * C * G);
  for (int c = 0; c < A; c++)
    for (int tr = 0; tr < T; tr++) {
      int from = m->from[tr], to = m->to[tr];
      for (int dp = 0; dp < G; dp++) {
        double num = st->transnum[ix_tr(m, s, c, tr, dp)];
        if (num != 0) {
          st->pop[ix_pop(m, s + 1, c, from, dp)] -= num;
          st->pop[ix_pop(m, s + 1, c, to, dp)] += num;
        }
      }
    }
}

// State::simulate(ti,tf) (src/state.cc:596-638).  pop_init[A][C][G] is used
// when ti == 0 (State::pop_init); otherwise pop[ti] must already be set.
// Returns 0, -4 on a NaN mean, -6 when a mean exceeds LARGE (utils.cc:150).
int or_simulate(const OrModel* m, const OrTables* tb, const double* pop_init, OrState* st, int ti, int tf) {
  const int A = m->A, T = m->T, G = m->G;
  if (ti == 0) std::memcpy(st->pop, pop_init, sizeof(double) * (size_t)A * m->C * G);
  for (int s = ti; s < tf; s++) {
    if (s == ti) or_set_I_from_pop(m, tb, st, s);
    else or_set_Imap_sett(m, tb, st, s);
    for (int c = 0; c < A; c++) {
      int rc = or_set_transmean(m, tb, st, s, c);
      if (rc) return rc;
      for (int tr = 0; tr < T; tr++)
        for (int dp = 0; dp < G; dp++) {
          size_t k = ix_tr(m, s, c, tr, dp);
          double mean = st->transmean[k];
          if (mean == 0) st->transnum[k] = 0;
          else if (m->stochastic) {
            int n = poisson_sample(mean);
            if (n < 0) return -6;
            st->transnum[k] = n;
          } else st->transnum[k] = mean;
        }
    }
    if (s < m->S - 1) or_update_pop(m, st, s);
  }
  return 0;
}

// Rebuilds pop, Imap/Idiag and transmean from a stored transnum trajectory:
// the arithmetic of State::initialise_from_particle (src/state.cc:294-322)
// and State::set_Imap (src/state.cc:127-160).
int or_replay(const OrModel* m, const OrTables* tb, const double* pop_init, OrState* st) {
  const int A = m->A, G = m->G;
  std::memcpy(st->pop, pop_init, sizeof(double) * (size_t)A * m->C * G);
  for (int s = 0; s < m->S; s++)
    if (s < m->S - 1) or_update_pop(m, st, s);
  for (int s = 0; s < m->S; s++) or_set_Imap_sett(m, tb, st, s);
  for (int s = 0; s < m->S; s++)
    for (int c = 0; c < A; c++) {
      int rc = or_set_transmean(m, tb, st, s, c);
      if (rc) return rc;
    }
  return 0;
}

// State::likelihood (src/state.cc:812-826)
double or_likelihood(const OrModel* m, const OrState* st) {
  double L = 0;
  size_t n = (size_t)m->S * m->A * m->T * m->G;
  for (size_t k = 0; k < n; k++) L += poisson_probability((int)st->transnum[k], st->transmean[k]);
  return L;
}

// Observation model.  Each observation i lists the state elements it sums,
// already in the reference's accumulation order (sett, area, [transitions
// then compartments], dp) of get_obs_value_section (src/obsmodel.cc:74-109):
//   el_kind 0 = transnum[s][c][x][dp], 1 = pop[s][c][x][dp].
// type: 0 normal (sd given), 2 Poisson, 3 negative binomial (shape);
// y == UNKNOWN contributes 0; y == THRESH uses thresh[i] (obs_prob 252-280).
struct OrObs {
  int nobs;
  const double* y;
  const double* sd;
  const int* type;
  const double* shape;
  const double* factor;
  const double* thresh;
  const int* el_ptr;   // [nobs+1]
  const int* el_kind;
  const int* el_s;
  const int* el_c;
  const int* el_x;
  const int* el_dp;
  const double* el_w;  // [nel] obs_spline value at the element's step, or null (no spline)
};

// get_obs_value_section restricted to steps [ti, tf).
void or_obs_values(const OrModel* m, const OrObs* ob, const OrState* st, int ti, int tf, double* value) {
  for (int i = 0; i < ob->nobs; i++) {
    double v = 0;
    for (int e = ob->el_ptr[i]; e < ob->el_ptr[i + 1]; e++) {
      int s = ob->el_s[e];
      if (s < ti || s >= tf) continue;
      double num = ob->el_kind[e] == 0 ? st->transnum[ix_tr(m, s, ob->el_c[e], ob->el_x[e], ob->el_dp[e])]
                                       : st->pop[ix_pop(m, s, ob->el_c[e], ob->el_x[e], ob->el_dp[e])];
      // With an observation spline: factor * num * spline(s) (src/obsmodel.cc:85-88).
      if (num != 0) v += ob->el_w ? ob->factor[i] * num * ob->el_w[e] : ob->factor[i] * num;
    }
    value[i] = v;
  }
}

// ObservationModel::obs_prob (src/obsmodel.cc:245-305)
double or_obs_prob(double value, double y, double sd, int type, double shape, double thresh) {
  if (value < 0) value = 0;
  if (y == UNKNOWN_OBS) return 0;
  if (y == THRESH_OBS) {
    switch (type) {
      case 0: case 1:
        return value < thresh ? normal_probability(thresh, thresh, sd * sd)
                              : normal_probability(thresh, value, sd * sd);
      case 2: {
        double lam = value < 1 ? 1 : value, sum = 0;
        for (unsigned k = 0; k <= thresh; k++) sum += std::exp(poisson_probability(k, lam));
        return std::log(sum);
      }
      case 3: {
        double mm = value < 1 ? 1 : value, sum = 0;
        for (unsigned k = 0; k <= thresh; k++) sum += std::exp(negative_binomial_probability(k, mm, shape));
        return std::log(sum);
      }
    }
    return NAN;
  }
  switch (type) {
    case 0: case 1: return normal_probability(y, value, sd * sd);
    case 2: { double lam = value < 1 ? 1 : value; return poisson_probability(int(y + 0.5), lam); }
    case 3: { double mm = value < 1 ? 1 : value; return negative_binomial_probability(int(y + 0.5), mm, shape); }
  }
  return NAN;
}

// ObservationModel::calculate / calculate_section: sum of obs_prob over the
// observations listed in sel (all observations when sel == NULL), with the
// values accumulated over steps [ti, tf).
double or_obs_loglik(const OrModel* m, const OrObs* ob, const OrState* st, int ti, int tf,
                     const int* sel, int nsel, double* value_out) {
  std::vector<double> value(ob->nobs);
  or_obs_values(m, ob, st, ti, tf, value.data());
  if (value_out) std::memcpy(value_out, value.data(), sizeof(double) * ob->nobs);
  double L = 0;
  int n = sel ? nsel : ob->nobs;
  for (int k = 0; k < n; k++) {
    int i = sel ? sel[k] : k;
    L += or_obs_prob(value[i], ob->y[i], ob->sd[i], ob->type[i], ob->shape[i], ob->thresh[i]);
  }
  return L;
}

// Mbp::mbp (src/mbp.cc:252-330) after the inbounds check and set_param.
// simu[S][A]: 1 = SIMU (fresh Poisson draw), 0 = MBP coupling.
// inf_update: 0 = INF_UPDATE (Imap recomputed by set_Imap_sett),
//             1 = INF_DIF_UPDATE (Imap_p = Imap_i + dImap).
int or_mbp(const OrModel* m, const OrTables* tbp, const double* pop_init, const OrState* ini, OrState* pro,
           const unsigned char* simu, int inf_update) {
  const int A = m->A, G = m->G, T = m->T;
  const size_t AG = (size_t)A * G;
  std::vector<double> dImap(AG, 0.0), dIdiag(AG, 0.0), dtn((size_t)A * T * G, 0.0);
  std::memcpy(pro->pop, pop_init, sizeof(double) * AG * m->C);       // mbp_initialise -> pop_init
  for (int s = 0; s < m->S; s++) {
    if (inf_update == 0) or_set_Imap_sett(m, tbp, pro, s);
    else {
      for (size_t v = 0; v < AG; v++) {                                  // set_Imap_using_dI
        pro->Imap[s * AG + v] = ini->Imap[s * AG + v] + dImap[v];
        pro->Idiag[s * AG + v] = ini->Idiag[s * AG + v] + dIdiag[v];
      }
    }
    for (int c = 0; c < A; c++) {
      int rc = or_set_transmean(m, tbp, pro, s, c);
      if (rc) return rc;
      int sorm_simu = simu ? simu[(size_t)s * A + c] : 0;
      for (int tr = 0; tr < T; tr++)
        for (int dp = 0; dp < G; dp++) {
          size_t k = ix_tr(m, s, c, tr, dp);
          int num_i = (int)ini->transnum[k], num_p = 0;
          double rate_p = pro->transmean[k];
          if (rate_p == 0) num_p = 0;
          else if (sorm_simu) num_p = poisson_sample(rate_p);
          else {
            double rate_i = ini->transmean[k];
            if (rate_p > rate_i) num_p = num_i + poisson_sample(rate_p - rate_i);
            else if (rate_p == rate_i) num_p = num_i;
            else num_p = (int)binomial_sample(rate_p / rate_i, (unsigned)num_i);
          }
          dtn[((size_t)c * T + tr) * G + dp] = num_p - num_i;
          pro->transnum[k] = num_p;
        }
    }
    if (s < m->S - 1) {
      or_update_pop(m, pro, s);
      if (inf_update == 1) update_I_from_transnum(m, tbp, dImap.data(), dIdiag.data(), dtn.data());
    }
  }
  return 0;
}

// Mbp::get_al (src/mbp.cc:334-361).  mode 0 = CUTOFF, 1 = INVT.
double or_get_al(int mode, double EFp, double EFi, double Prp, double Pri, double invT, double EFcut) {
  double al = 0;
  if (mode == 0) { if (EFp < EFcut) al = std::exp(Prp - Pri); }
  else al = std::exp(-invT * 0.5 * (EFp - EFi) + Prp - Pri);
  if (al > 1) al = 1;
  return al;
}

// Sample streams for pinning against the reference's utils.cc.
void or_poisson_stream(unsigned seed, double lam, int n, int* out) {
  g_gen.seed(seed);
  for (int i = 0; i < n; i++) out[i] = poisson_sample(lam);
}
void or_binomial_stream(unsigned seed, double p, unsigned nn, int n, int* out) {
  g_gen.seed(seed);
  for (int i = 0; i < n; i++) out[i] = (int)binomial_sample(p, nn);
}
double or_normal_probability(double x, double mean, double var) { return normal_probability(x, mean, var); }
double or_poisson_probability(int i, double lam) { return poisson_probability(i, lam); }
double or_negative_binomial_probability(unsigned v, double m, double shape) {
  return negative_binomial_probability(v, m, shape);
}

}  // extern "C"
