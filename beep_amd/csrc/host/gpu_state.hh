// C++ host mirror of the reference's State / ObservationModel / Mbp interface
// on top of the C ABI (include/beep_gpu.h).  A driver written against the
// reference's classes (src/state.hh:11-133, src/obsmodel.hh:8-39,
// src/mbp.hh:15-82) uses these with the same method names, argument meaning
// and results; the arithmetic runs on the GPU.
//
// Differences a caller sees:
//  - random numbers come from counter-based Philox streams keyed by
//    (seed, stream, instance key, step, cell, transition), not from the
//    process-global libstdc++ engine; each simulate / mbp call of a state
//    advances that state's stream counter;
//  - errors are returned as exceptions (beep::Error carrying the BEEP_* code)
//    instead of emsg()'s exit(), so a caller can choose to exit.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "beep_gpu.h"

namespace beep {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// src/struct.hh:176-188 (transnum flattened [S][A][T][G]).
struct Particle {
  std::vector<double> paramval;
  std::vector<int32_t> transnum;
  double EF = 0;
  unsigned run = 0;
};

enum ObsModelMode { CUTOFF = 0, INVT = 1 };       // src/consts.hh ObsModelMode
enum InfUpdate { INF_UPDATE = 0, INF_DIF_UPDATE = 1 };
enum Status { SUCCESS = 0, FAIL = 1 };

// What Details / Data / Model / ObservationModel hold for one model, plus the
// device context with room for `capacity` states.
class GpuModel {
 public:
  GpuModel(const beep_spec& spec, int capacity, int keep = BEEP_KEEP_ALL, int device = 0, size_t hbm_budget = 0,
           std::vector<std::string> param_names = {});
  ~GpuModel();
  GpuModel(const GpuModel&) = delete;
  GpuModel& operator=(const GpuModel&) = delete;

  beep_ctx* ctx() const { return ctx_; }
  int A, G, C, T, S, nparam, nobs, ndatatable, capacity, steps_per_day, nsection;
  std::vector<std::string> param_names;
  uint64_t seed = 1;
  void check(int rc) const;

 private:
  beep_ctx* ctx_ = nullptr;
};

// One latent state (src/state.hh): owns instance `inst` and table slot `slot`.
class GpuState {
 public:
  GpuState(GpuModel& model, int inst, uint32_t key);

  double EF = 0;                      // -2 log(observation model)
  double Pr = 0;                      // log prior
  std::vector<double> paramval;

  void set_param(const std::vector<double>& paramv);   // State::set_param (src/state.cc:66-91)
  void simulate(const std::vector<double>& paramv);    // src/state.cc:582-592
  void simulate(unsigned ti, unsigned tf);             // src/state.cc:596-638
  void set_EF();                                       // src/state.cc:95-98
  void set_Pr();                                       // src/state.cc:102-105
  double likelihood() const;                           // src/state.cc:812-826
  void initialise_from_particle(const Particle& part); // src/state.cc:294-322
  Particle create_particle(unsigned run) const;        // src/state.cc:326-335
  void save(const std::string& file) const;            // src/state.cc:709-739
  void load(const std::string& file, unsigned ndivision_post);   // src/state.cc:743-808

  std::vector<int32_t> transnum() const;               // [S][A][T][G]
  std::vector<float> transmean() const;                // [S][A][T][G]
  std::vector<int32_t> pop() const;                    // [S][A][C][G]

  int inst() const { return inst_; }
  int slot() const { return inst_; }
  uint32_t key() const { return key_; }
  GpuModel& model() const { return m_; }
  uint32_t next_stream() { return stream_++ & 0xFFFF; }

 private:
  GpuModel& m_;
  int inst_;
  uint32_t key_;
  uint32_t stream_ = 0;
  bool inbounds_ = true;
};

// src/obsmodel.hh: the observation model of a GpuModel.
class GpuObservationModel {
 public:
  explicit GpuObservationModel(GpuModel& m) : m_(m) {}
  double calculate(const GpuState* s) const;                        // src/obsmodel.cc:22-35
  double calculate_section(const GpuState* s, unsigned sec) const;  // src/obsmodel.cc:39-58
  std::vector<double> get_EF_datatable(const GpuState* s) const;    // src/obsmodel.cc:309-322
  std::vector<double> get_obs_value(const GpuState* s) const;       // src/obsmodel.cc:62-70

 private:
  GpuModel& m_;
};

// src/mbp.hh: model-based proposals between two states (initial / propose),
// swapped on acceptance like Mbp::swap_initial_propose_state.
class GpuMbp {
 public:
  GpuMbp(ObsModelMode mode, GpuModel& m, GpuState& a, GpuState& b) : mode_(mode), m_(m), initial_(&a), propose_(&b) {}
  double invT = 1.0;
  double EFcut = 0.0;
  // Mbp::mbp (src/mbp.cc:252-330); simu_mask [S][A] (1 = simulate afresh) may be empty.
  Status mbp(const std::vector<double>& paramv, InfUpdate inf_update,
             const std::vector<uint8_t>& simu_mask = std::vector<uint8_t>());
  // Mbp::get_al (src/mbp.cc:334-361): sets the proposal's EF and Pr.
  double get_al();
  // MVN::MH + swap (src/mvn.cc:233-249, src/mbp.cc:243-248): accepts with
  // probability al (drawn on the device) and swaps initial / propose.
  bool accept(double dlogq_extra = 0.0);
  GpuState& initial() { return *initial_; }
  GpuState& propose() { return *propose_; }

 private:
  ObsModelMode mode_;
  GpuModel& m_;
  GpuState* initial_;
  GpuState* propose_;
  uint32_t stream_ = 0x8000;
};

// PMCMC::obs_prob and Mpi::particle_sample (src/pmcmc.cc:180-268,
// src/mpi.cc:215-385) for N particles held in instances inst[0..N): the
// whole filter (sections, bootstrap resampling, state moves) runs on the
// device in one call.
class GpuParticleFilter {
 public:
  GpuParticleFilter(GpuModel& m, std::vector<int32_t> inst, int slot, uint32_t key0 = 0);
  double invT = 1.0;                                   // PMCMC::invT
  double obs_prob(const std::vector<double>& paramv);  // src/pmcmc.cc:180-207
  // The trajectory traced back from the first particle of the last
  // bootstrap, gathered into instance dst and returned as a Particle.
  Particle particle_sample(unsigned run, int dst);     // src/mpi.cc:331-385
  const std::vector<int32_t>& backpart() const { return backpart_; }   // [(nsection+1)][N]

 private:
  GpuModel& m_;
  std::vector<int32_t> inst_;
  std::vector<uint32_t> key_;
  int slot_;
  uint32_t stream_ = 0;
  std::vector<double> paramval_;
  double obprob_ = 0;
  std::vector<int32_t> backpart_;
};

}  // namespace beep
