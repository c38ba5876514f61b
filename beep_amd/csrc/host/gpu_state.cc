// C++ host mirror of State / ObservationModel / Mbp over the C ABI.
#include "gpu_state.hh"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace beep {

namespace {
std::vector<std::string> split_csv(const std::string& line) {
  std::vector<std::string> out;
  std::string cur;
  std::stringstream ss(line);
  while (std::getline(ss, cur, ',')) {
    while (!cur.empty() && (cur.back() == '\r' || cur.back() == ' ')) cur.pop_back();
    while (!cur.empty() && cur.front() == ' ') cur.erase(cur.begin());
    out.push_back(cur);
  }
  return out;
}
}  // namespace

GpuModel::GpuModel(const beep_spec& spec, int cap, int keep, int device, size_t hbm_budget,
                   std::vector<std::string> names)
    : param_names(std::move(names)) {
  int rc = beep_create(&spec, device, hbm_budget, cap, cap, keep, &ctx_);
  if (rc != BEEP_OK) {
    std::string msg = ctx_ ? beep_last_error(ctx_) : "beep_create failed";
    if (ctx_) beep_destroy(ctx_);
    ctx_ = nullptr;
    throw Error(rc, msg);
  }
  int32_t d[10];
  beep_get_dims(ctx_, d);
  A = d[0]; G = d[1]; C = d[2]; T = d[3]; S = d[4]; nparam = d[5]; nobs = d[6]; ndatatable = d[7];
  capacity = d[8];
  steps_per_day = spec.steps_per_day;
  nsection = spec.nsection;
  if ((int)param_names.size() != nparam) {
    param_names.clear();
    for (int i = 0; i < nparam; i++) param_names.push_back("param" + std::to_string(i));
  }
}

GpuModel::~GpuModel() {
  if (ctx_) beep_destroy(ctx_);
}

void GpuModel::check(int rc) const {
  if (rc != BEEP_OK) throw Error(rc, beep_last_error(ctx_));
}

// ---- GpuState ----

GpuState::GpuState(GpuModel& m, int inst, uint32_t key) : m_(m), inst_(inst), key_(key) {
  if (inst < 0 || inst >= m.capacity) throw Error(BEEP_EINVAL, "instance out of range");
}

void GpuState::set_param(const std::vector<double>& pv) {
  if ((int)pv.size() != m_.nparam) throw Error(BEEP_EINVAL, "wrong number of parameters");
  paramval = pv;
  int32_t slot = inst_, inb = 1;
  double pr = 0;
  m_.check(beep_set_param(m_.ctx(), 1, &slot, pv.data(), &pr, &inb));
  Pr = pr;
  inbounds_ = inb != 0;
}

void GpuState::simulate(const std::vector<double>& pv) {
  set_param(pv);
  simulate(0, m_.S);
  set_EF();
  set_Pr();
}

void GpuState::simulate(unsigned ti, unsigned tf) {
  if (ti == 0) stream_++;               // a fresh run draws a fresh stream; sections continue it
  int32_t inst = inst_, slot = inst_;
  uint32_t key = key_;
  m_.check(beep_simulate(m_.ctx(), 1, &inst, &slot, &key, (int)ti, (int)tf, m_.seed, stream_ & 0xFFFF, 0));
}

void GpuState::set_EF() {
  int32_t inst = inst_;
  double L = 0;
  m_.check(beep_obs_loglik(m_.ctx(), 1, &inst, -1, &L, nullptr));
  EF = -2 * L;
}

void GpuState::set_Pr() {
  // Model::prior was evaluated by set_param (src/model.cc:1590-1659).
}

double GpuState::likelihood() const {
  int32_t inst = inst_;
  double L = 0;
  m_.check(beep_pois_loglik(m_.ctx(), 1, &inst, &L));
  return L;
}

void GpuState::initialise_from_particle(const Particle& part) {
  if ((int)part.transnum.size() != m_.S * m_.A * m_.T * m_.G) throw Error(BEEP_EINVAL, "particle transnum size");
  set_param(part.paramval);
  EF = part.EF;
  m_.check(beep_put_transnum(m_.ctx(), inst_, inst_, part.transnum.data()));
}

Particle GpuState::create_particle(unsigned run) const {
  Particle p;
  p.run = run;
  p.EF = EF;
  p.paramval = paramval;
  p.transnum = transnum();
  return p;
}

std::vector<int32_t> GpuState::transnum() const {
  std::vector<int32_t> out((size_t)m_.S * m_.A * m_.T * m_.G);
  m_.check(beep_get_transnum(m_.ctx(), inst_, out.data()));
  return out;
}

std::vector<float> GpuState::transmean() const {
  std::vector<float> out((size_t)m_.S * m_.A * m_.T * m_.G);
  m_.check(beep_get_transmean(m_.ctx(), inst_, out.data()));
  return out;
}

std::vector<int32_t> GpuState::pop() const {
  std::vector<int32_t> out((size_t)m_.S * m_.A * m_.C * m_.G);
  m_.check(beep_get_pop(m_.ctx(), inst_, out.data()));
  return out;
}

// The posterior sample files of State::save (src/state.cc:709-739).
void GpuState::save(const std::string& file) const {
  std::ofstream outp(file + "_parameter.csv");
  if (!outp) throw Error(BEEP_EINVAL, "Cannot open the file '" + file + "_parameter.csv'");
  outp.precision(17);
  outp << "Parameter,Value\n";
  outp << "Error function," << EF << "\n";
  for (int th = 0; th < m_.nparam; th++) outp << m_.param_names[th] << "," << paramval[th] << "\n";
  std::ofstream outt(file + "_transition.csv");
  if (!outt) throw Error(BEEP_EINVAL, "Cannot open the file '" + file + "_transition.csv'");
  outt << "Time,Transitions (ordered by area x trans x democatpos)\n";
  std::vector<int32_t> tn = transnum();
  size_t k = 0;
  for (int s = 0; s < m_.S; s++) {
    outt << (s + 0.5) / m_.steps_per_day;           // details.division_time (src/details.cc:102-106)
    for (int i = 0; i < m_.A * m_.T * m_.G; i++) outt << "," << tn[k++];
    outt << "\n";
  }
}

// State::load (src/state.cc:743-808): steps from ndivision_post on are zero.
void GpuState::load(const std::string& file, unsigned ndivision_post) {
  if ((int)ndivision_post > m_.S) throw Error(BEEP_EINVAL, "ndivision_post beyond the horizon");
  Particle part;
  std::ifstream inp(file + "_parameter.csv");
  if (!inp) throw Error(BEEP_EINVAL, "Cannot open the file '" + file + "_parameter.csv'");
  std::string line;
  std::getline(inp, line);
  std::getline(inp, line);
  auto spl = split_csv(line);
  if (spl.size() != 2) throw Error(BEEP_EINVAL, "Problem loading the file '" + file + "_parameter.csv'");
  part.EF = std::stod(spl[1]);
  part.paramval.resize(m_.nparam);
  for (int th = 0; th < m_.nparam; th++) {
    std::getline(inp, line);
    spl = split_csv(line);
    if (spl.size() != 2 || spl[0] != m_.param_names[th])
      throw Error(BEEP_EINVAL, "Problem loading the file '" + file + "_parameter.csv'");
    part.paramval[th] = std::stod(spl[1]);
  }
  std::ifstream int_(file + "_transition.csv");
  if (!int_) throw Error(BEEP_EINVAL, "Cannot open the file '" + file + "_transition.csv'");
  std::getline(int_, line);
  const size_t row = (size_t)m_.A * m_.T * m_.G;
  part.transnum.assign((size_t)m_.S * row, 0);
  for (unsigned s = 0; s < ndivision_post; s++) {
    if (!std::getline(int_, line)) throw Error(BEEP_EINVAL, "Problem loading the file '" + file + "_transition.csv'");
    spl = split_csv(line);
    if (spl.size() != row + 1) throw Error(BEEP_EINVAL, "Problem loading the file '" + file + "_transition.csv'");
    for (size_t i = 0; i < row; i++) part.transnum[s * row + i] = (int32_t)std::lround(std::stod(spl[i + 1]));
  }
  initialise_from_particle(part);
}

// ---- GpuObservationModel ----

double GpuObservationModel::calculate(const GpuState* s) const {
  int32_t inst = s->inst();
  double L = 0;
  m_.check(beep_obs_loglik(m_.ctx(), 1, &inst, -1, &L, nullptr));
  return L;
}

double GpuObservationModel::calculate_section(const GpuState* s, unsigned sec) const {
  int32_t inst = s->inst();
  double L = 0;
  m_.check(beep_obs_loglik(m_.ctx(), 1, &inst, (int)sec, &L, nullptr));
  return L;
}

std::vector<double> GpuObservationModel::get_EF_datatable(const GpuState* s) const {
  int32_t inst = s->inst();
  double L = 0;
  std::vector<double> out(m_.ndatatable);
  m_.check(beep_obs_loglik(m_.ctx(), 1, &inst, -1, &L, out.data()));
  return out;
}

std::vector<double> GpuObservationModel::get_obs_value(const GpuState* s) const {
  std::vector<double> out(m_.nobs);
  m_.check(beep_obs_values(m_.ctx(), s->inst(), out.data()));
  return out;
}

// ---- GpuMbp ----

Status GpuMbp::mbp(const std::vector<double>& paramv, InfUpdate inf_update, const std::vector<uint8_t>& simu_mask) {
  propose_->set_param(paramv);
  int32_t init = initial_->inst(), prop = propose_->inst(), slot = propose_->slot();
  uint32_t key = propose_->key();
  // Model::inbounds is evaluated host-side by set_param; out of bounds -> FAIL.
  double pr = 0;
  int32_t inb = 1;
  m_.check(beep_set_param(m_.ctx(), 1, &slot, paramv.data(), &pr, &inb));
  if (!inb) return FAIL;
  if (!simu_mask.empty() && simu_mask.size() != (size_t)m_.S * m_.A) throw Error(BEEP_EINVAL, "simu mask size");
  m_.check(beep_mbp(m_.ctx(), 1, &init, &prop, &slot, &key, simu_mask.empty() ? nullptr : simu_mask.data(), m_.seed,
                    propose_->next_stream() | 0x4000, (int)inf_update));
  return SUCCESS;
}

double GpuMbp::get_al() {
  propose_->set_EF();
  propose_->set_Pr();
  int32_t init = initial_->inst(), prop = propose_->inst();
  uint32_t key = propose_->key();
  double dlogq = propose_->Pr - initial_->Pr, al = 0;
  uint8_t acc = 0;
  m_.check(beep_accept(m_.ctx(), 1, &init, &prop, (int)mode_, invT, EFcut, &dlogq, &key, m_.seed, stream_ & 0xFFFF,
                       &al, &acc));
  return al;
}

bool GpuMbp::accept(double dlogq_extra) {
  int32_t init = initial_->inst(), prop = propose_->inst();
  uint32_t key = propose_->key();
  double dlogq = propose_->Pr - initial_->Pr + dlogq_extra, al = 0;
  uint8_t acc = 0;
  m_.check(beep_accept(m_.ctx(), 1, &init, &prop, (int)mode_, invT, EFcut, &dlogq, &key, m_.seed, stream_ & 0xFFFF,
                       &al, &acc));
  stream_++;
  if (acc) std::swap(initial_, propose_);
  return acc != 0;
}

// ---- GpuParticleFilter ----

GpuParticleFilter::GpuParticleFilter(GpuModel& m, std::vector<int32_t> inst, int slot, uint32_t key0)
    : m_(m), inst_(std::move(inst)), slot_(slot) {
  if (inst_.empty()) throw Error(BEEP_EINVAL, "no particles");
  for (size_t i = 0; i < inst_.size(); i++) key_.push_back(key0 + (uint32_t)i);
}

double GpuParticleFilter::obs_prob(const std::vector<double>& paramv) {
  if ((int)paramv.size() != m_.nparam) throw Error(BEEP_EINVAL, "wrong number of parameters");
  int32_t slot = slot_, inb = 1;
  double pr = 0;
  m_.check(beep_set_param(m_.ctx(), 1, &slot, paramv.data(), &pr, &inb));   // every particle's set_param
  paramval_ = paramv;
  const int nsec = m_.nsection;
  backpart_.assign((size_t)(nsec + 1) * inst_.size(), 0);
  m_.check(beep_pmcmc_obs_prob(m_.ctx(), (int)inst_.size(), inst_.data(), slot_, key_.data(), invT, m_.seed,
                               stream_++ & 0xFFFF, &obprob_, backpart_.data()));
  return obprob_;
}

Particle GpuParticleFilter::particle_sample(unsigned run, int dst) {
  m_.check(beep_pmcmc_particle_sample(m_.ctx(), dst));
  Particle p;
  p.run = run;
  p.paramval = paramval_;
  p.transnum.resize((size_t)m_.S * m_.A * m_.T * m_.G);
  m_.check(beep_get_transnum(m_.ctx(), dst, p.transnum.data()));
  return p;
}

}  // namespace beep
