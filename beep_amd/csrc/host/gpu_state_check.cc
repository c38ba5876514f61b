// Self-check of the C++ host mirror (GpuModel / GpuState / GpuObservationModel
// / GpuMbp) on the GPU: the single-area SEIR with daily I->R observations.
// Usage: gpu_state_check <output directory>
// Writes <dir>/sample0_parameter.csv and <dir>/sample0_transition.csv (the
// State::save format) and prints one JSON line; exits non-zero on a failure.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "gpu_state.hh"

using namespace beep;

static int failures = 0;
#define CHECK(cond, what)                                         \
  do {                                                            \
    if (!(cond)) {                                                \
      std::fprintf(stderr, "check failed: %s\n", what);           \
      failures++;                                                 \
    }                                                             \
  } while (0)

int main(int argc, char** argv) {
  std::string dir = argc > 1 ? argv[1] : ".";
  const int period = 100;
  // Parameters: R (uniform 1..4), mE = 5, mI = 7, infI = 1 (fixed).
  beep_prior prior[4] = {{BEEP_PRIOR_UNIFORM, 1.0, 4.0}, {BEEP_PRIOR_FIXED, 5.0, 0}, {BEEP_PRIOR_FIXED, 7.0, 0},
                         {BEEP_PRIOR_FIXED, 1.0, 0}};
  int32_t mE[1] = {1}, mI[1] = {2};
  beep_comp comp[4] = {{-1, nullptr, 1}, {-1, mE, 1}, {3, mI, 1}, {-1, nullptr, 1}};
  beep_trans trans[3] = {{0, 1, 1, nullptr}, {1, 2, 0, nullptr}, {2, 3, 0, nullptr}};
  double kt[2] = {0, (double)period}, mf[2] = {1, 1};
  int32_t kp[2] = {0, 0};
  beep_spline spl = {2, kt, kp, mf, -1};
  int32_t pop[4] = {99990, 0, 10, 0};
  double agedist[1] = {1.0};
  std::vector<beep_obs> obs(period);
  int32_t area0[1] = {0}, dp0[1] = {0}, trIR[1] = {2};
  for (int d = 0; d < period; d++) {
    double y = std::floor(600.0 * std::exp(-0.5 * std::pow((d - 45.0) / 12.0, 2)) + 0.5);
    beep_obs& o = obs[d];
    o = beep_obs{};
    o.y = y;
    o.sd = std::max(5.0, 0.1 * y);
    o.factor = 1.0;
    o.type = BEEP_OBS_NORMAL;
    o.sett_i = d;
    o.sett_f = d + 1;
    o.narea = 1; o.area = area0;
    o.ndp = 1; o.dp = dp0;
    o.ntr = 1; o.tr = trIR;
  }
  int32_t sec_ti[2] = {0, 50}, sec_tf[2] = {50, 100};
  beep_spec spec{};
  spec.A = 1; spec.G = 1; spec.C = 4; spec.T = 3;
  spec.period = period; spec.steps_per_day = 1;
  spec.nparam = 4; spec.prior = prior; spec.comp = comp; spec.trans = trans;
  spec.nspline = 1; spec.spline = &spl; spec.R_spline = 0; spec.Rfactor_param = -1;
  spec.efoi_spline = -1; spec.efoi_agedist = agedist; spec.geo_spline = -1;
  spec.pop_init = pop;
  spec.nobs = period; spec.obs = obs.data();
  spec.nsection = 2; spec.section_ti = sec_ti; spec.section_tf = sec_tf;

  try {
    GpuModel m(spec, 4 + 32, BEEP_KEEP_ALL, 0, 0, {"R", "mE", "mI", "infI"});
    GpuObservationModel om(m);
    GpuState a(m, 0, 11), b(m, 1, 12), c(m, 2, 13);
    std::vector<double> theta = {2.1, 5.0, 7.0, 1.0};
    a.simulate(theta);
    CHECK(std::fabs(a.EF + 2 * om.calculate(&a)) <= 1e-9 * std::fabs(a.EF), "EF == -2 calculate");
    CHECK(std::fabs(a.Pr - std::log(1.0 / 3.0)) < 1e-14, "Pr == log(1/(4-1))");
    double Ls = om.calculate_section(&a, 0) + om.calculate_section(&a, 1);
    CHECK(std::fabs(Ls + a.EF / 2) <= 1e-9 * std::fabs(a.EF), "sections sum to the whole");
    auto dt = om.get_EF_datatable(&a);
    CHECK(dt.size() == 1 && std::fabs(dt[0] - a.EF) <= 1e-9 * std::fabs(a.EF), "EF by data table");
    auto pa = a.pop();
    for (int s = 0; s < m.S; s++) {
      long tot = 0;
      for (int co = 0; co < 4; co++) tot += pa[(size_t)s * 4 + co];
      CHECK(tot == 100000, "population conserved");
    }
    // initialise_from_particle reproduces the state.
    Particle part = a.create_particle(0);
    b.initialise_from_particle(part);
    CHECK(b.transnum() == a.transnum() && b.pop() == a.pop(), "initialise_from_particle: transnum and pop");
    CHECK(b.transmean() == a.transmean(), "initialise_from_particle: transmean");
    CHECK(b.likelihood() == a.likelihood(), "initialise_from_particle: likelihood");
    b.set_EF();
    CHECK(b.EF == a.EF, "initialise_from_particle: EF");
    // save / load round trip.
    a.save(dir + "/sample0");
    c.load(dir + "/sample0", m.S);
    CHECK(c.transnum() == a.transnum() && c.paramval == a.paramval && c.EF == a.EF, "save/load round trip");
    c.load(dir + "/sample0", 60);
    auto tc = c.transnum();
    bool zero_after = true;
    for (size_t k = 60 * 3; k < tc.size(); k++) zero_after &= tc[k] == 0;
    CHECK(zero_after, "load(ndivision_post) zeroes later steps");
    // MBPs: identical parameters reproduce the trajectory and al == 1.
    GpuMbp mbp(INVT, m, a, b);
    CHECK(mbp.mbp(theta, INF_DIF_UPDATE) == SUCCESS, "mbp success");
    CHECK(b.transnum() == a.transnum(), "mbp identity");
    double al = mbp.get_al();
    CHECK(std::fabs(al - 1.0) < 1e-12, "al == 1 for the identity proposal");
    std::vector<double> out_of_prior = {5.0, 5.0, 7.0, 1.0};
    CHECK(mbp.mbp(out_of_prior, INF_DIF_UPDATE) == FAIL, "mbp outside the prior fails");
    std::vector<double> theta2 = {2.4, 5.0, 7.0, 1.0};
    CHECK(mbp.mbp(theta2, INF_DIF_UPDATE) == SUCCESS, "mbp to a new R");
    double al2 = mbp.get_al();
    bool acc = mbp.accept();
    // Particle filter over the two sections with 32 particles (instances 4..35).
    std::vector<int32_t> pinst;
    for (int i = 0; i < 32; i++) pinst.push_back(4 + i);
    GpuParticleFilter pf(m, pinst, 3, 500);
    pf.invT = 1.0;
    double ob = pf.obs_prob(theta);
    CHECK(std::isfinite(ob), "particle filter estimate is finite");
    const auto& bpk = pf.backpart();
    CHECK(bpk.size() == 3 * 32, "backpart rows");
    bool kept = true;
    for (int k = 1; k <= 2; k++)
      for (int p = 0; p < 32; p++) kept &= bpk[k * 32 + bpk[k * 32 + p]] == bpk[k * 32 + p];
    CHECK(kept, "every ancestor keeps its own place");
    Particle ps = pf.particle_sample(0, 3);
    long tot_inf = 0;
    for (int s = 0; s < m.S; s++) tot_inf += ps.transnum[(size_t)s * 3 + 0];
    CHECK(tot_inf > 0 && ps.paramval == theta, "sampled particle");
    std::printf("{\"EF\": %.17g, \"Pr\": %.17g, \"likelihood\": %.17g, \"al_identity\": %.17g, \"al2\": %.17g, "
                "\"accepted\": %d, \"pf_obprob\": %.17g, \"failures\": %d}\n",
                a.EF, a.Pr, a.likelihood(), al, al2, acc ? 1 : 0, ob, failures);
  } catch (const Error& e) {
    std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
    return 2;
  }
  return failures ? 1 : 0;
}
