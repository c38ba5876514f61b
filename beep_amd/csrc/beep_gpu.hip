// C ABI of the MI355X hot path: contexts, parameter slots, instances, and the
// launches of the simulation / observation / likelihood / MBP / accept kernels.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <map>
#include <mutex>
#include <tuple>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "beep_gpu.h"
#include "host_model.hh"
#include "launch_geometry.hh"
#include "kernels.hh"
#include "obs_kernels.hh"
#include "pmcmc_kernels.hh"
#include "step_kernels.hh"
// kJitDeviceRng, kJitFusedCommon, kJitFusedSpec, kJitStepSpec: the texts of
// device_rng.hh, fused_common.hh, fused_spec.hh and step_spec.hh (generated
// by build.py).
#include "jit_sources.inc"

using beepdev::Job;
using beepdev::TopoDev;

// The kernels specialised for one compartment structure (fused_spec.hh,
// step_spec.hh), compiled once per (device, structure) and kept for the life
// of the process.
struct SpecKernels {
  hipModule_t mod = nullptr;
  hipFunction_t fn[4] = {};        // fused kernel, indexed by MODE_SIM / MBP / REPLAY / ANY
  hipFunction_t step[3] = {};      // step-wise transition kernel, MODE_SIM / MBP / REPLAY
  int max_dyn_lds = 65536;         // dynamic LDS the fused kernels accept (the device's, less static use)
};

struct beep_ctx {
  beephost::HostModel m;
  int device = 0;
  hipStream_t stream = nullptr;
  int keep = 0, capacity = 0, nslot = 0;
  size_t budget = 0, allocated = 0;
  std::string err;
  std::vector<void*> allocs;

  // topology
  TopoDev* d_topo = nullptr;
  TopoDev h_topo;
  // parameter slots
  float* d_tab = nullptr;
  size_t slot_stride = 0;
  std::vector<double> slot_param;   // [nslot][nparam] last parameters of each slot
  std::vector<char> slot_set;
  // instances
  int* d_state = nullptr;
  int* d_tn = nullptr;
  float* d_tm = nullptr;
  int* d_pt = nullptr;
  double* d_EF = nullptr;           // last EF of each instance
  int* d_err = nullptr;
  size_t state_sz = 0, traj_sz = 0, pop_traj_sz = 0;
  // observations
  beepdev::ObsDev obs;
  beepdev::ObsDev* d_obs = nullptr;
  std::vector<int> sec_ptr;
  int* d_sec_ptr = nullptr;
  int* d_sec_obs = nullptr;
  // staging
  Job* d_jobs = nullptr;
  size_t jobs_cap = 0;
  Job* d_bjobs = nullptr;          // fused jobs of mixed batches launched from this context
  size_t bjobs_cap = 0;
  void* d_scratch = nullptr;
  size_t scratch_cap = 0;
  // timing of the last launch of each kernel kind
  hipEvent_t ev[6] = {};
  double timing[4] = {0, 0, 0, 0};
  // step-wise path (populations in HBM, one launch per kernel per step)
  int stepwise = 0, dense = 0, Apad = 0;
  float* d_M = nullptr;            // [Apad][Apad] dense off-diagonal mixing, transposed, fragment order
  float* d_D = nullptr;            // [capacity*G padded to 64][Apad], fragment order
  float* d_I = nullptr;            // [capacity*G][Apad]
  uint64_t* d_rngp = nullptr;      // seed, stream read by the captured step kernels
  struct Graph {
    int mode = -1, n = 0, t0 = 0, t1 = 0;
    const Job* jobs = nullptr;
    long long rows = 0, cells_y = 0;   // launch geometry captured
    const void* spec = nullptr;        // kernels captured (specialised or generic)
    hipGraphExec_t exec = nullptr;
  };
  std::vector<Graph> graphs;       // captured step loops, most recently made last
  // particle filter: ancestry rows [(nsection+1)][pf_n] then the instances [pf_n]
  int* d_sec_ti = nullptr;
  int* d_sec_tf = nullptr;
  int* d_pf = nullptr;
  size_t pf_cap = 0;
  int pf_n = 0;
  int pf_slot = -1;                // the parameter slot of the last particle filter run
  // structure-specialised fused kernel (null: generic kernel or step-wise path)
  SpecKernels* spec = nullptr;
  size_t spec_lds = 0;
  // Streams over which a mixed batch launched from this context spreads its
  // size classes (created on first use), and the event they start from.
  hipStream_t bstream[4] = {};
  hipEvent_t bstart = nullptr;
  hipEvent_t bend[4] = {};         // end of each batch stream's work
  bool obs_general = false;        // observations other than plain normal ones
  bool obs_pop = false;            // observations counting population elements
  // Observation splines: their values per slot, [nslot][nospl][S] (FP64), and
  // the slot each instance was last simulated, proposed or replayed with.
  int nospl = 0;
  double* d_ospl = nullptr;
  std::vector<int> inst_slot;
  int* d_inst_slot = nullptr;
  bool inst_slot_dirty = false;
  // Pinned host staging of beep_set_param: the slot blocks derived in
  // parallel, uploaded with one copy.
  float* h_stage = nullptr;
  size_t h_stage_cap = 0;          // floats
  double* h_stage64 = nullptr;
  size_t h_stage64_cap = 0;        // doubles
};

// Device memory of the allocation-free entry point (beep_simulate_noalloc):
// the job descriptors of one call, at most BEEP_NOALLOC_BYTES per device.
// Each device holds its own copy until the process ends; host code reaches it
// only through the address hipGetSymbolAddress gives on that device
// (g_noalloc_ptr), never through the array's name.
#define BEEP_NOALLOC_BYTES (1u << 20)
static __device__ __align__(256) unsigned char scratch[BEEP_NOALLOC_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};

int fail(beep_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(c, x)                                                              \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) return fail(c, BEEP_EHIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

int dalloc(beep_ctx* c, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return BEEP_OK;
  if (c->budget && c->allocated + bytes > c->budget)
    return fail(c, BEEP_ENOMEM, "HBM budget exceeded: %zu + %zu > %zu bytes", c->allocated, bytes, c->budget);
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(c, BEEP_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  c->allocated += bytes;
  c->allocs.push_back(*p);
  return BEEP_OK;
}

template <class T>
int upload(beep_ctx* c, T** d, const std::vector<T>& h) {
  size_t bytes = std::max<size_t>(1, h.size()) * sizeof(T);
  int rc = dalloc(c, (void**)d, bytes);
  if (rc) return rc;
  if (!h.empty()) HIPCHK(c, hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return BEEP_OK;
}

std::vector<float> to_f32(const std::vector<double>& v) { return std::vector<float>(v.begin(), v.end()); }

int ensure_jobs(beep_ctx* c, size_t n) {
  if (n <= c->jobs_cap) return BEEP_OK;
  if (c->d_jobs) {
    for (auto& g : c->graphs)     // captured graphs refer to the old job buffer
      if (g.jobs >= c->d_jobs && g.jobs < c->d_jobs + c->jobs_cap) {
        HIPCHK(c, hipGraphExecDestroy(g.exec));
        g.exec = nullptr;
      }
    c->graphs.erase(std::remove_if(c->graphs.begin(), c->graphs.end(), [](const beep_ctx::Graph& g) { return !g.exec; }),
                    c->graphs.end());
    HIPCHK(c, hipFree(c->d_jobs));
    c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), (void*)c->d_jobs));
    c->allocated -= c->jobs_cap * sizeof(Job);
    c->d_jobs = nullptr;
    c->jobs_cap = 0;
  }
  size_t cap = std::max<size_t>(n, 1024);
  int rc = dalloc(c, (void**)&c->d_jobs, cap * sizeof(Job));
  if (rc) return rc;
  c->jobs_cap = cap;
  return BEEP_OK;
}

int ensure_scratch(beep_ctx* c, size_t bytes) {
  if (bytes <= c->scratch_cap) return BEEP_OK;
  if (c->d_scratch) {
    HIPCHK(c, hipFree(c->d_scratch));
    c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), c->d_scratch));
    c->allocated -= c->scratch_cap;
    c->d_scratch = nullptr;
    c->scratch_cap = 0;
  }
  size_t cap = std::max<size_t>(bytes, 1 << 16);
  int rc = dalloc(c, &c->d_scratch, cap);
  if (rc) return rc;
  c->scratch_cap = cap;
  return BEEP_OK;
}

void record(beep_ctx* c, int k) { (void)hipEventRecord(c->ev[k], c->stream); }
double elapsed(beep_ctx* c, int k0, int k1) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, c->ev[k0], c->ev[k1]) != hipSuccess) return -1;
  return ms;
}

bool valid_inst(const beep_ctx* c, int i) { return i >= 0 && i < c->capacity; }
bool valid_slot(const beep_ctx* c, int s) { return s >= 0 && s < c->nslot; }

// The instances one call writes must be in range and distinct: each is one
// trajectory and one state, and two jobs writing the same one would race
// (and, on the step-wise path, index infectivity rows past the context's
// capacity).  `read`, when given, lists instances the call only reads (the
// initial states of MBPs): none of them may be written by the same call.
int check_instances(beep_ctx* c, int n, const int32_t* written, const int32_t* read, const char* what) {
  std::vector<char> seen(c->capacity, 0);
  for (int i = 0; i < n; i++) {
    if (!valid_inst(c, written[i])) return fail(c, BEEP_EINVAL, "%s: instance %d out of range", what, written[i]);
    if (seen[written[i]]++) return fail(c, BEEP_EINVAL, "%s: instance %d written twice", what, written[i]);
  }
  for (int i = 0; read && i < n; i++) {
    if (!valid_inst(c, read[i])) return fail(c, BEEP_EINVAL, "%s: instance %d out of range", what, read[i]);
    if (seen[read[i]]) return fail(c, BEEP_EINVAL, "%s: instance %d is both read and written", what, read[i]);
  }
  return BEEP_OK;
}

int check_device_error(beep_ctx* c) {
  int e = 0;
  HIPCHK(c, hipMemcpyAsync(&e, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (e) {
    HIPCHK(c, hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (e & 4) return fail(c, BEEP_ENAN, "PMCMC bootstrap: non-finite likelihood estimate or weights");
    if (e & 1) return fail(c, BEEP_ENAN, "NaN transition mean");
    return fail(c, BEEP_ELARGE, "transition mean above 1e9");
  }
  return BEEP_OK;
}

// Once per device, with that device current (called from beep_create): the
// device address of `scratch` on it.
int noalloc_setup(beep_ctx* c) {
  if (c->device < 0 || c->device >= kMaxDevices) return fail(c, BEEP_EINVAL, "device %d out of range", c->device);
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr[c->device]) return BEEP_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(c, BEEP_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", c->device, hipGetErrorString(e));
  g_noalloc_ptr[c->device] = p;
  return BEEP_OK;
}

// Threads per workgroup for the fused kernel: whole wavefronts covering the
// cells of one instance in one pass, at most BEEP_FUSED_MAX_BLOCK.
int fused_block(int AG) {
  int b = (AG + 63) / 64 * 64;
  return b > BEEP_FUSED_MAX_BLOCK ? BEEP_FUSED_MAX_BLOCK : b;
}

// Threads per workgroup for the specialised fused kernel: 1, 2, 4 or 8
// wavefronts, so that its register-limited resident wavefronts per CU are
// whole workgroups (measured on 100 areas x 4 age groups: 512 threads beat
// the one-pass 448 by 5%); 16 for instances of more than 1024 cells, whose
// LDS leaves room for one workgroup per CU, so that the CU still has sixteen
// wavefronts to switch between.
int spec_block(int AG) {
  int w = (AG + 63) / 64;
  w = w <= 1 ? 1 : w <= 2 ? 2 : w <= 4 ? 4 : w <= 16 ? 8 : 16;
  return 64 * w;
}

template <int MODE>
int launch_fused_t(beep_ctx* c, const Job* djobs, int njobs, size_t lds, int block, uint64_t seed, uint32_t stream,
                   hipStream_t st = nullptr) {
  if (njobs == 0) return BEEP_OK;
  {
    std::string e = beephost::fused_geometry(njobs, block, lds, 160 * 1024, BEEP_FUSED_MAX_BLOCK);
    if (!e.empty()) return fail(c, BEEP_EINVAL, "%s", e.c_str());
  }
  auto kfn = beepdev::k_fused<MODE>;
  if (lds > 65536) HIPCHK(c, hipFuncSetAttribute((const void*)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kfn, dim3(njobs), dim3(block), lds, st ? st : c->stream, djobs, seed, stream);
  HIPCHK(c, hipGetLastError());
  return BEEP_OK;
}

std::mutex g_jit_mu;
std::map<std::string, SpecKernels*> g_jit;

// The constants fused_spec.hh is compiled with.
// Minimum resident wavefronts per SIMD the specialised kernels are compiled
// for (it caps their registers): measured on 100 areas x 4 age groups with 8
// transitions, 6 beat 4 (no spills either way) and 8 spilled; fewer
// transitions need fewer registers.  BEEP_SPEC_WAVES / BEEP_STEP_WAVES
// override.
int spec_waves(const beephost::HostModel& m, const char* env, int dflt) {
  if (const char* e = std::getenv(env)) return std::max(1, std::min(8, std::atoi(e)));
  return dflt;
}

std::string spec_prelude(const beephost::HostModel& m) {
  // At least 4: a workgroup of BEEP_SPEC_MAX_BLOCK threads is four wavefronts
  // per SIMD.
  const int waves = std::max(4, spec_waves(m, "BEEP_SPEC_WAVES", m.T <= 4 ? 8 : m.T <= 8 ? 6 : 4));
  const int step_waves = spec_waves(m, "BEEP_STEP_WAVES", 4);
  std::ostringstream o;
  o << "namespace beepspec {\n";
  o << "constexpr int C = " << m.C << ", T = " << m.T << ", G = " << m.G << ", CP = " << (m.C + 3) / 4 * 4
    << ", INF_TR = " << m.inf_tr << ", MIN_WAVES = " << waves << ", STEP_WAVES = " << step_waves
    << ", NSTRAIN = " << m.nstrain << ", DPM = " << m.G / m.nstrain << ";\n";
  o << "constexpr int FROM[" << m.T << "] = {";
  for (int t = 0; t < m.T; t++) o << (t ? ", " : "") << m.trans[t].from;
  o << "};\nconstexpr int TO[" << m.T << "] = {";
  for (int t = 0; t < m.T; t++) o << (t ? ", " : "") << m.trans[t].to;
  o << "};\nconstexpr bool INFC[" << m.C << "] = {";
  for (int c = 0; c < m.C; c++) o << (c ? ", " : "") << (m.comp[c].inf_param >= 0 ? "true" : "false");
  o << "};\n}  // namespace beepspec\n";
  return o.str();
}

int spec_kernels(beep_ctx* c, SpecKernels** out) {
  const std::string prelude = spec_prelude(c->m);
  const std::string key = std::to_string(c->device) + "\n" + prelude;
  std::lock_guard<std::mutex> lk(g_jit_mu);
  auto it = g_jit.find(key);
  if (it != g_jit.end()) {
    *out = it->second;
    return BEEP_OK;
  }
  const std::string src =
      prelude + "#include \"fused_common.hh\"\n#include \"fused_spec.hh\"\n#include \"step_spec.hh\"\n";
  const char* hdr[4] = {kJitDeviceRng, kJitFusedCommon, kJitFusedSpec, kJitStepSpec};
  const char* names[4] = {"device_rng.hh", "fused_common.hh", "fused_spec.hh", "step_spec.hh"};
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "beep_spec.hip", 4, hdr, names) != HIPRTC_SUCCESS)
    return fail(c, BEEP_EHIP, "hiprtcCreateProgram failed");
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=fast",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  hiprtcResult r = hiprtcCompileProgram(prog, 8, opts);
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n + 1, '\0');
    hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    return fail(c, BEEP_EHIP, "compiling the specialised kernel: %s: %.400s", hiprtcGetErrorString(r), log.c_str());
  }
  size_t n = 0;
  hiprtcGetCodeSize(prog, &n);
  std::vector<char> code(n);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  auto* k = new SpecKernels();
  hipError_t e = hipModuleLoadData(&k->mod, code.data());
  const char* fn[4] = {"beep_fused_spec_sim", "beep_fused_spec_mbp", "beep_fused_spec_replay", "beep_fused_spec_any"};
  for (int i = 0; e == hipSuccess && i < 4; i++) e = hipModuleGetFunction(&k->fn[i], k->mod, fn[i]);
  const char* sn[3] = {"beep_step_spec_sim", "beep_step_spec_mbp", "beep_step_spec_replay"};
  for (int i = 0; e == hipSuccess && i < 3; i++) e = hipModuleGetFunction(&k->step[i], k->mod, sn[i]);
  if (e != hipSuccess) {
    delete k;
    return fail(c, BEEP_EHIP, "loading the specialised kernel: %s", hipGetErrorString(e));
  }
  int lim = 0;
  for (int i = 0; i < 4; i++) {
    int v = 0;
    if (hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_MAX_DYNAMIC_SHARED_SIZE_BYTES, k->fn[i]) == hipSuccess && v > 0)
      lim = i == 0 ? v : std::min(lim, v);
  }
  if (lim > 0) k->max_dyn_lds = lim;
  g_jit[key] = k;
  *out = k;
  return BEEP_OK;
}

int launch_spec(beep_ctx* c, const SpecKernels* k, int mode, const Job* djobs, int njobs, size_t lds, int block,
                uint64_t seed, uint32_t stream, hipStream_t st = nullptr) {
  if (njobs == 0) return BEEP_OK;
  {
    std::string e = beephost::fused_geometry(njobs, block, lds, (size_t)k->max_dyn_lds, BEEP_SPEC_MAX_BLOCK);
    if (!e.empty()) return fail(c, BEEP_EINVAL, "%s", e.c_str());
  }
  void* args[3] = {(void*)&djobs, (void*)&seed, (void*)&stream};
  HIPCHK(c, hipModuleLaunchKernel(k->fn[mode], (unsigned)njobs, 1, 1, (unsigned)block, 1, 1, (unsigned)lds,
                                  st ? st : c->stream, args, nullptr));
  return BEEP_OK;
}

// The step-wise path over [t0, t1) for the n jobs at djobs: initial
// infectivity, then per step the contraction (MFMA GEMM when dense) and the
// transition kernel.  The launch sequence is captured once into a hipGraph and
// replayed while (mode, n, t0, t1, jobs) stay the same; seed and stream are
// read from device memory so one graph serves every call.
int launch_stepwise(beep_ctx* c, int mode, const Job* djobs, int n, int t0, int t1, uint64_t seed, uint32_t stream) {
  const auto& m = c->m;
  beephost::StepGeometry geo;
  {
    std::string e = beephost::stepwise_geometry(m.A, m.G, c->Apad, n, c->capacity, BEEP_STEP_BLOCK, &geo);
    if (!e.empty()) return fail(c, BEEP_EINVAL, "%s", e.c_str());
  }
  // The job descriptors must lie inside the context's job buffer: the graph
  // reads them there at every replay.
  if (!c->d_jobs || djobs < c->d_jobs || djobs + n > c->d_jobs + c->jobs_cap)
    return fail(c, BEEP_EINVAL, "step-wise launch: job descriptors outside the context's job buffer");
  uint64_t rp[2] = {seed, (uint64_t)stream};
  HIPCHK(c, hipMemcpyAsync(c->d_rngp, rp, sizeof rp, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (const auto& g : c->graphs)
    if (g.mode == mode && g.n == n && g.t0 == t0 && g.t1 == t1 && g.jobs == djobs) {
      // A cached graph is replayed only with the geometry it was captured with.
      if (g.rows != geo.rows || g.cells_y != geo.cell.y || g.spec != c->spec)
        return fail(c, BEEP_EINVAL, "step-wise launch: cached graph does not match the launch geometry");
      HIPCHK(c, hipGraphLaunch(g.exec, c->stream));
      return BEEP_OK;
    }
  if (c->graphs.size() >= 32) {
    // The oldest graph may still be running on the stream: it is destroyed
    // only after the stream has drained.
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipGraphExecDestroy(c->graphs.front().exec));
    c->graphs.erase(c->graphs.begin());
  }
  beep_ctx::Graph G;
  const int rows = (int)geo.rows;
  const dim3 gcell((unsigned)geo.cell.x, (unsigned)geo.cell.y);
  hipGraph_t graph;
  (void)hipGetLastError();               // a stale error of an earlier call is not this loop's
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(beepdev::k_step_init, gcell, dim3(BEEP_STEP_BLOCK), 0, c->stream, djobs, c->d_D, c->Apad);
  for (int s = t0; s < t1; s++) {
    if (c->dense) {
      beepdev::DenseMix p{c->d_D, c->d_M, c->d_I, c->h_topo.diag, c->h_topo.fac, djobs, rows, m.A, c->Apad, m.G, s,
                          c->h_topo.off_geo};
      hipLaunchKernelGGL(beepdev::k_dense_contract, dim3((unsigned)geo.dense.x, (unsigned)geo.dense.y), dim3(256), 0,
                         c->stream, p);
    } else {
      beepdev::SparseMix p{c->d_D, c->d_I, c->h_topo.diag, c->h_topo.fac, c->h_topo.in_ptr, c->h_topo.in_src,
                           c->h_topo.in_w, djobs, rows, m.A, c->Apad, m.G, s, c->h_topo.off_geo};
      hipLaunchKernelGGL(beepdev::k_sparse_contract, dim3((unsigned)geo.sparse.x, (unsigned)geo.sparse.y), dim3(256),
                         0, c->stream, p);
    }
    if (c->spec) {
      int Apad = c->Apad, ss = s;
      void* args[6] = {(void*)&djobs, (void*)&c->d_I, (void*)&c->d_D, (void*)&Apad, (void*)&ss, (void*)&c->d_rngp};
      hipError_t le = hipModuleLaunchKernel(c->spec->step[mode], gcell.x, gcell.y, 1, BEEP_STEP_BLOCK, 1, 1, 0,
                                            c->stream, args, nullptr);
      if (le != hipSuccess) {
        hipGraph_t g2 = nullptr;
        (void)hipStreamEndCapture(c->stream, &g2);
        if (g2) (void)hipGraphDestroy(g2);
        return fail(c, BEEP_EHIP, "step kernel launch under capture: %s", hipGetErrorString(le));
      }
    } else if (mode == beepdev::MODE_SIM)
      hipLaunchKernelGGL(beepdev::k_step_trans<beepdev::MODE_SIM>, gcell, dim3(BEEP_STEP_BLOCK), 0, c->stream, djobs,
                         c->d_I, c->d_D, c->Apad, s, c->d_rngp);
    else if (mode == beepdev::MODE_MBP)
      hipLaunchKernelGGL(beepdev::k_step_trans<beepdev::MODE_MBP>, gcell, dim3(BEEP_STEP_BLOCK), 0, c->stream, djobs,
                         c->d_I, c->d_D, c->Apad, s, c->d_rngp);
    else
      hipLaunchKernelGGL(beepdev::k_step_trans<beepdev::MODE_REPLAY>, gcell, dim3(BEEP_STEP_BLOCK), 0, c->stream, djobs,
                         c->d_I, c->d_D, c->Apad, s, c->d_rngp);
    // A launch refused while capturing (resources, configuration) is an
    // error of the call, not a silently empty step.
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
      hipGraph_t g2 = nullptr;
      (void)hipStreamEndCapture(c->stream, &g2);
      if (g2) (void)hipGraphDestroy(g2);
      return fail(c, BEEP_EHIP, "step %d launch under capture: %s", s, hipGetErrorString(le));
    }
  }
  hipError_t e = hipStreamEndCapture(c->stream, &graph);
  if (e != hipSuccess) return fail(c, BEEP_EHIP, "graph capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(c, BEEP_EHIP, "graph instantiate: %s", hipGetErrorString(e));
  G.mode = mode; G.n = n; G.t0 = t0; G.t1 = t1; G.jobs = djobs;
  G.rows = geo.rows; G.cells_y = geo.cell.y; G.spec = c->spec;
  c->graphs.push_back(G);
  HIPCHK(c, hipGraphLaunch(G.exec, c->stream));
  return BEEP_OK;
}

int launch_fused(beep_ctx* c, int mode, const Job* djobs, int njobs, uint64_t seed, uint32_t stream) {
  const auto& m = c->m;
  if (c->spec && !c->stepwise)
    return launch_spec(c, c->spec, mode, djobs, njobs, c->spec_lds, spec_block(m.AG), seed, stream);
  size_t lds = beepdev::fused_lds_bytes(m.A, m.G, m.C, m.T, c->h_topo.nnz, c->h_topo.stage_mix);
  if (lds > 160 * 1024) return fail(c, BEEP_EINVAL, "instance needs %zu bytes of LDS (max 163840)", lds);
  int block = fused_block(m.AG);
  if (mode == beepdev::MODE_ANY) return launch_fused_t<beepdev::MODE_ANY>(c, djobs, njobs, lds, block, seed, stream);
  if (mode == beepdev::MODE_SIM) return launch_fused_t<beepdev::MODE_SIM>(c, djobs, njobs, lds, block, seed, stream);
  if (mode == beepdev::MODE_MBP) return launch_fused_t<beepdev::MODE_MBP>(c, djobs, njobs, lds, block, seed, stream);
  return launch_fused_t<beepdev::MODE_REPLAY>(c, djobs, njobs, lds, block, seed, stream);
}

// The observation log-likelihood of n instances over the observations sel
// (all when null), into out[b * (1 + ndatatable)].
void launch_obs_loglik(beep_ctx* c, int n, const int* d_inst, const int* sel, int nsel, double* out, int store_EF,
                       bool want_dt, hipStream_t st = nullptr) {
  using K = decltype(&beepdev::k_obs_loglik<true, true, true, false>);
  // With observation splines only the general variants are used.
  static const K ks[12] = {
      beepdev::k_obs_loglik<false, false, false, false>, beepdev::k_obs_loglik<false, false, true, false>,
      beepdev::k_obs_loglik<false, true, false, false>,  beepdev::k_obs_loglik<false, true, true, false>,
      beepdev::k_obs_loglik<true, false, false, false>,  beepdev::k_obs_loglik<true, false, true, false>,
      beepdev::k_obs_loglik<true, true, false, false>,   beepdev::k_obs_loglik<true, true, true, false>,
      beepdev::k_obs_loglik<false, true, false, true>,   beepdev::k_obs_loglik<false, true, true, true>,
      beepdev::k_obs_loglik<true, true, false, true>,    beepdev::k_obs_loglik<true, true, true, true>};
  K k = c->nospl ? ks[8 + (want_dt ? 2 : 0) + (c->obs_pop ? 1 : 0)]
                 : ks[(want_dt ? 4 : 0) + (c->obs_general ? 2 : 0) + (c->obs_pop ? 1 : 0)];
  const int ni = want_dt ? beepdev::OBS_INSTANCES<true> : beepdev::OBS_INSTANCES<false>;
  // Per-table sums: one pass per eight data tables.
  const int npass = want_dt ? (c->m.ndatatable + 7) / 8 : 1;
  for (int p = 0; p < npass; p++)
    hipLaunchKernelGGL(k, dim3((n + ni - 1) / ni), dim3(256), 0, st ? st : c->stream, c->d_obs, n, d_inst, c->d_tn,
                       c->d_pt, c->traj_sz, c->pop_traj_sz, sel, nsel, out, c->d_EF, store_EF, c->d_ospl,
                       c->d_inst_slot, 8 * p);
}

// Before an observation kernel reads inst_slot: the host copy, uploaded when
// it changed (a synchronous copy: later jobs change the host copy).
int sync_inst_slot(beep_ctx* c) {
  if (!c->nospl || !c->inst_slot_dirty) return BEEP_OK;
  HIPCHK(c, hipMemcpyAsync(c->d_inst_slot, c->inst_slot.data(), c->inst_slot.size() * sizeof(int),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->inst_slot_dirty = false;
  return BEEP_OK;
}

// Defined with the mixed batches, which share it.
int obs_loglik_chunked(beep_ctx* c0, const std::vector<beep_ctx*>& ctxs, const std::vector<const int32_t*>& inst,
                       const std::vector<int>& ninst, bool store_EF, double* out, int ev0);

Job make_job(beep_ctx* c, int inst, int slot, uint32_t key, int t0, int t1) {
  Job j;
  std::memset(&j, 0, sizeof j);
  j.topo = c->d_topo;
  j.tab = c->d_tab + (size_t)slot * c->slot_stride;
  j.state = c->d_state + (size_t)inst * c->state_sz;
  j.tn = c->d_tn ? c->d_tn + (size_t)inst * c->traj_sz : nullptr;
  j.tm = c->d_tm ? c->d_tm + (size_t)inst * c->traj_sz : nullptr;
  j.pt = c->d_pt ? c->d_pt + (size_t)inst * c->pop_traj_sz : nullptr;
  j.err = c->d_err;
  j.key = key;
  j.t0 = t0;
  j.t1 = t1;
  j.from_init = t0 == 0 ? 1 : 0;
  j.mode = beepdev::MODE_SIM;
  c->inst_slot[inst] = slot;
  c->inst_slot_dirty = true;
  return j;
}

// Copies host-side job descriptors to the device (the copy is ordered before
// the kernels queued after it on the context stream).
int put_jobs(beep_ctx* c, const std::vector<Job>& jobs) {
  int rc = ensure_jobs(c, jobs.size());
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->d_jobs, jobs.data(), jobs.size() * sizeof(Job), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));   // the host vector is released after return
  return BEEP_OK;
}

template <class F>
int debug_run(int device, size_t in_bytes, const void* in, size_t in2_bytes, const void* in2, size_t out_bytes, void* out,
              F launch) {
  if (hipSetDevice(device) != hipSuccess) return BEEP_EHIP;
  void *d_in = nullptr, *d_in2 = nullptr, *d_out = nullptr;
  int rc = BEEP_OK;
  if (hipMalloc(&d_in, in_bytes) != hipSuccess || hipMalloc(&d_out, out_bytes) != hipSuccess ||
      (in2_bytes && hipMalloc(&d_in2, in2_bytes) != hipSuccess))
    rc = BEEP_ENOMEM;
  if (!rc && hipMemcpy(d_in, in, in_bytes, hipMemcpyHostToDevice) != hipSuccess) rc = BEEP_EHIP;
  if (!rc && in2_bytes && hipMemcpy(d_in2, in2, in2_bytes, hipMemcpyHostToDevice) != hipSuccess) rc = BEEP_EHIP;
  if (!rc) {
    launch(d_in, d_in2, d_out);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = BEEP_EHIP;
  }
  if (!rc && hipMemcpy(out, d_out, out_bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = BEEP_EHIP;
  if (d_in) (void)hipFree(d_in);
  if (d_in2) (void)hipFree(d_in2);
  if (d_out) (void)hipFree(d_out);
  return rc;
}

}  // namespace

extern "C" {

const char* beep_last_error(const beep_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int beep_spec_tables(const beep_spec* spec, const double* paramval, double* rate, double* sus, double* inf,
                     double* Ntime, double* beta_af, double* efoi, double* geo, double* Pr, int32_t* inb, char* err,
                     int errlen) {
  auto put_err = [&](const std::string& e) {
    if (err && errlen > 0) { std::strncpy(err, e.c_str(), errlen - 1); err[errlen - 1] = 0; }
  };
  if (!spec || !paramval) return BEEP_EINVAL;
  beephost::HostModel m;
  std::string e = m.build(*spec);
  if (!e.empty()) { put_err(e); return BEEP_EINVAL; }
  beephost::Tables tb;
  int rc = m.derive(paramval, tb, e);
  if (rc) { put_err(e); return rc; }
  if (rate) std::copy(tb.rate.begin(), tb.rate.end(), rate);
  if (sus) std::copy(tb.sus.begin(), tb.sus.end(), sus);
  if (inf) std::copy(tb.inf.begin(), tb.inf.end(), inf);
  if (Ntime) std::copy(tb.Ntime.begin(), tb.Ntime.end(), Ntime);
  if (beta_af) {
    std::vector<double> ba = m.beta_af(tb);
    std::copy(ba.begin(), ba.end(), beta_af);
  }
  if (efoi) {
    std::vector<double> eg = m.efoi_cells(tb);
    std::copy(eg.begin(), eg.end(), efoi);
  }
  if (geo) std::copy(tb.geo.begin(), tb.geo.end(), geo);
  if (Pr) *Pr = m.prior_logpdf(paramval);
  if (inb) *inb = m.inbounds(paramval) ? 1 : 0;
  return BEEP_OK;
}

int beep_create(const beep_spec* spec, int device, size_t hbm_budget, int capacity, int nslot, int keep,
                beep_ctx** out) {
  if (!out) return BEEP_EINVAL;
  *out = nullptr;
  if (!spec || capacity <= 0 || nslot <= 0) return BEEP_EINVAL;
  beep_ctx* c = new beep_ctx();
  std::string e = c->m.build(*spec);
  if (!e.empty()) {
    c->err = e;
    *out = c;   // the caller reads the message, then destroys the context
    return BEEP_EINVAL;
  }
  c->device = device;
  c->budget = hbm_budget;
  c->capacity = capacity;
  c->nslot = nslot;
  c->keep = keep;
  *out = c;
  HIPCHK(c, hipSetDevice(device));
  int rc0 = noalloc_setup(c);
  if (rc0) return rc0;
  HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (auto& e : c->ev) HIPCHK(c, hipEventCreate(&e));
  const auto& m = c->m;

  // Topology.
  int *d_from, *d_to, *d_in_ptr, *d_in_src, *d_pop_init, *d_order;
  float *d_diag, *d_fac, *d_in_w;
  std::vector<int> from(m.T), to(m.T);
  for (int t = 0; t < m.T; t++) { from[t] = m.trans[t].from; to[t] = m.trans[t].to; }
  std::vector<int> pop_dev((size_t)m.C * m.AG);     // [C][AG]
  for (int a = 0; a < m.A; a++)
    for (int co = 0; co < m.C; co++)
      for (int g = 0; g < m.G; g++)
        pop_dev[(size_t)co * m.AG + a * m.G + g] = m.pop_init[((size_t)a * m.C + co) * m.G + g];
  // Areas by number of incoming mixing entries, most first (ties by index):
  // the order in which the fused kernels' lanes take areas for the mixing.
  std::vector<int> order(m.A);
  for (int a = 0; a < m.A; a++) order[a] = a;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return m.in_ptr[x + 1] - m.in_ptr[x] > m.in_ptr[y + 1] - m.in_ptr[y];
  });
  int rc;
  if ((rc = upload(c, &d_from, from)) || (rc = upload(c, &d_to, to)) || (rc = upload(c, &d_in_ptr, m.in_ptr)) ||
      (rc = upload(c, &d_in_src, m.in_src)) || (rc = upload(c, &d_pop_init, pop_dev)) ||
      (rc = upload(c, &d_order, order)))
    return rc;
  std::vector<float> diag = to_f32(m.M_diag), fac = to_f32(m.M_factor), inw = to_f32(m.in_w);
  if ((rc = upload(c, &d_diag, diag)) || (rc = upload(c, &d_fac, fac)) || (rc = upload(c, &d_in_w, inw))) return rc;

  TopoDev& tp = c->h_topo;
  tp.A = m.A; tp.G = m.G; tp.C = m.C; tp.T = m.T; tp.S = m.S; tp.AG = m.AG; tp.inf_tr = m.inf_tr; tp.spd = m.spd;
  tp.nnz = (int)m.in_src.size();
  // The mixing CSR is staged in LDS when small, and only if the instance still
  // fits the fused kernel's 160 KB with it.
  tp.stage_mix = beepdev::fused_mix_bytes(m.A, tp.nnz) <= 32768 &&
                 beepdev::fused_lds_bytes(m.A, m.G, m.C, m.T, tp.nnz, 1) <= 160 * 1024;
  tp.dt = (float)m.dt;
  int off = 0;
  tp.off_rate = off; off += m.T * m.G;
  tp.off_sus = off; off += m.G;
  tp.off_inf = off; off += m.C;
  tp.N_tv = m.agemod.empty() ? 0 : 1;
  tp.off_N = off; off += m.G * m.G * (tp.N_tv ? m.S : 1);
  tp.off_efoi = off; off += m.S * m.nE;
  tp.off_geo = off; off += m.S;
  tp.off_beta = off; off += m.S * m.nR;
  // The area factor: in each slot when it depends on the parameters, else one
  // table for the context.
  {
    float* d_agedist = nullptr;
    if ((rc = upload(c, &d_agedist, to_f32(m.egrp_agedist)))) return rc;
    tp.agedist = d_agedist;
    int *d_rgrp = nullptr, *d_egrp = nullptr;
    if ((rc = upload(c, &d_rgrp, m.rgrp_area)) || (rc = upload(c, &d_egrp, m.egrp_cell))) return rc;
    tp.nR = m.nR;
    tp.nE = m.nE;
    tp.rgrp = d_rgrp;
    tp.egrp = d_egrp;
  }
  // Prediction-mode modifications: context-wide multiplier tables, null when
  // the model has none of a kind (the kernels then read nothing more).
  tp.mod_beta = tp.mod_efoi = tp.mod_trans = nullptr;
  {
    const std::vector<double>* mt[3] = {&m.mod_beta, &m.mod_efoi, &m.mod_trans};
    const float** dst[3] = {&tp.mod_beta, &tp.mod_efoi, &tp.mod_trans};
    for (int k = 0; k < 3; k++)
      if (!mt[k]->empty()) {
        float* d = nullptr;
        if ((rc = upload(c, &d, to_f32(*mt[k])))) return rc;
        *dst[k] = d;
      }
  }
  tp.off_af = -1;
  tp.af = nullptr;
  if (m.af_param) {
    tp.off_af = off;
    off += m.period * m.A;
  } else {
    float* d_af = nullptr;
    if ((rc = upload(c, &d_af, to_f32(m.af_const)))) return rc;
    tp.af = d_af;
  }
  c->slot_stride = ((size_t)off + 63) / 64 * 64;
  tp.from = d_from; tp.to = d_to; tp.diag = d_diag; tp.fac = d_fac;
  tp.in_ptr = d_in_ptr; tp.in_src = d_in_src; tp.in_w = d_in_w; tp.mix_order = d_order; tp.pop_init = d_pop_init;

  // Path: the fused per-instance kernel when an instance fits in LDS and the
  // mixing is sparse; otherwise the step-wise path, with the contraction on
  // the matrix cores when the mixing is dense.
  if (m.C > 16) return fail(c, BEEP_EINVAL, "at most 16 compartments are supported");
  // Dense: a quarter of the matrix filled and rows of 32 entries or more on
  // average -- where the per-cell gather would cost more than a batched GEMM.
  // k-nearest-neighbour mixing of a few dozen areas stays on the fused path.
  c->dense = (size_t)tp.nnz * 4 >= (size_t)m.A * m.A && (size_t)tp.nnz >= (size_t)32 * m.A && m.A >= 32;
  size_t lds = beepdev::fused_lds_bytes(m.A, m.G, m.C, m.T, tp.nnz, tp.stage_mix);
  c->stepwise = (keep & BEEP_STEPWISE) || c->dense || lds > 160 * 1024;
  if (!(keep & BEEP_GENERIC_KERNEL) && m.C <= 16 && m.T <= 64) {
    // The step-wise transition kernel's LDS: populations, queue, tables.
    size_t step_lds = 4 * ((size_t)16 * BEEP_STEP_BLOCK + 2 * (size_t)m.T * BEEP_STEP_BLOCK + 1 + (size_t)m.T * m.G +
                           (size_t)m.G * m.G + m.G + m.C);
    // The specialised kernel stages the mixing CSR in LDS whenever it fits:
    // on the 100-area benchmark model, gathering from LDS at 3 resident
    // instances per CU beat gathering through the caches at 4.
    // BEEP_STAGE_MIX=0/1 forces the choice (a tuning knob).
    int stage = tp.stage_mix;
    const int grouped = m.nR > 1 || m.nE > 1;
    if (!c->stepwise) {
      const bool small = beepdev::fused_mix_bytes(m.A, tp.nnz) <= 32768;
      stage = small && beepdev::spec_lds_bytes(m.A, m.G, m.C, m.T, tp.nnz, 1, m.nstrain, grouped, m.nE) <= 160 * 1024;
      if (const char* e = std::getenv("BEEP_STAGE_MIX")) stage = small && std::atoi(e) != 0;
    }
    size_t sl = beepdev::spec_lds_bytes(m.A, m.G, m.C, m.T, tp.nnz, stage, m.nstrain, grouped, m.nE);
    if (c->stepwise ? step_lds <= 65536 : sl <= 160 * 1024) {
      SpecKernels* k = nullptr;
      if ((rc = spec_kernels(c, &k))) return rc;
      if (c->stepwise || sl <= (size_t)k->max_dyn_lds) {
        c->spec = k;
        c->spec_lds = c->stepwise ? 0 : sl;
        if (!c->stepwise) tp.stage_mix = stage;
      }
    }
  }
  // Strains pool susceptibles across cells, which only the specialised fused
  // kernel does.
  if (m.nstrain > 1 && (c->stepwise || !c->spec))
    return fail(c, BEEP_EINVAL, "a model with strains runs on the specialised fused kernel only (an instance in LDS, "
                                "sparse mixing, not BEEP_GENERIC_KERNEL or BEEP_STEPWISE)");
  if ((m.nR > 1 || m.nE > 1) && !c->spec)
    return fail(c, BEEP_EINVAL, "R or external-force splines per area or strain run on the specialised kernels only "
                                "(not BEEP_GENERIC_KERNEL)");
  if ((rc = dalloc(c, (void**)&c->d_topo, sizeof(TopoDev)))) return rc;
  HIPCHK(c, hipMemcpy(c->d_topo, &tp, sizeof tp, hipMemcpyHostToDevice));
  c->Apad = (m.A + 63) / 64 * 64;
  if (c->dense) {
    // The GEMM's right operand: M transposed, in fragment order (Mt[j][a] = M[a][j]).
    std::vector<float> M((size_t)c->Apad * c->Apad, 0.f);
    for (int a = 0; a < m.A; a++)
      for (int k = m.M_rowptr[a]; k < m.M_rowptr[a + 1]; k++)
        M[beepdev::frag_at(m.M_to[k], a, c->Apad)] += (float)m.M_val[k];
    if ((rc = upload(c, &c->d_M, M))) return rc;
  }
  if (c->stepwise) {
    // D in fragment order over rows padded to the GEMM's 64-row tiles (the
    // padding rows and columns stay zero); I row-major.
    const size_t rows_pad = ((size_t)capacity * m.G + 63) / 64 * 64;
    size_t szD = rows_pad * c->Apad * sizeof(float), sz = (size_t)capacity * m.G * c->Apad * sizeof(float);
    if ((rc = dalloc(c, (void**)&c->d_D, szD)) || (rc = dalloc(c, (void**)&c->d_I, sz))) return rc;
    HIPCHK(c, hipMemset(c->d_D, 0, szD));
    HIPCHK(c, hipMemset(c->d_I, 0, sz));
    if ((rc = dalloc(c, (void**)&c->d_rngp, 2 * sizeof(uint64_t)))) return rc;
  }

  // Slots and instances.
  if ((rc = dalloc(c, (void**)&c->d_tab, c->slot_stride * nslot * sizeof(float)))) return rc;
  HIPCHK(c, hipMemset(c->d_tab, 0, c->slot_stride * nslot * sizeof(float)));
  c->slot_param.assign((size_t)nslot * m.nparam, 0.0);
  c->slot_set.assign(nslot, 0);
  c->inst_slot.assign(capacity, -1);
  c->nospl = (int)m.obs_splines.size();
  if (c->nospl) {
    size_t sz = (size_t)nslot * c->nospl * m.S * sizeof(double);
    if ((rc = dalloc(c, (void**)&c->d_ospl, sz))) return rc;
    HIPCHK(c, hipMemset(c->d_ospl, 0, sz));
    if ((rc = dalloc(c, (void**)&c->d_inst_slot, capacity * sizeof(int)))) return rc;
    c->inst_slot_dirty = true;
  }
  c->state_sz = (size_t)m.C * m.AG;
  c->traj_sz = (size_t)m.S * m.T * m.AG;
  c->pop_traj_sz = (size_t)m.S * m.C * m.AG;
  if ((rc = dalloc(c, (void**)&c->d_state, c->state_sz * capacity * sizeof(int)))) return rc;
  // Trajectories start at zero, as the State's vectors do (steps a partial
  // simulation does not reach read as zero).
  if (keep & BEEP_KEEP_TRANSNUM) {
    if ((rc = dalloc(c, (void**)&c->d_tn, c->traj_sz * capacity * sizeof(int)))) return rc;
    HIPCHK(c, hipMemset(c->d_tn, 0, c->traj_sz * capacity * sizeof(int)));
  }
  if (keep & BEEP_KEEP_TRANSMEAN) {
    if ((rc = dalloc(c, (void**)&c->d_tm, c->traj_sz * capacity * sizeof(float)))) return rc;
    HIPCHK(c, hipMemset(c->d_tm, 0, c->traj_sz * capacity * sizeof(float)));
  }
  if (keep & BEEP_KEEP_POP) {
    if ((rc = dalloc(c, (void**)&c->d_pt, c->pop_traj_sz * capacity * sizeof(int)))) return rc;
    HIPCHK(c, hipMemset(c->d_pt, 0, c->pop_traj_sz * capacity * sizeof(int)));
  }
  if ((rc = dalloc(c, (void**)&c->d_EF, capacity * sizeof(double)))) return rc;
  HIPCHK(c, hipMemset(c->d_EF, 0, capacity * sizeof(double)));
  if ((rc = dalloc(c, (void**)&c->d_err, sizeof(int)))) return rc;
  HIPCHK(c, hipMemset(c->d_err, 0, sizeof(int)));
  for (int i = 0; i < capacity; i++)
    HIPCHK(c, hipMemcpy(c->d_state + (size_t)i * c->state_sz, d_pop_init, c->state_sz * sizeof(int),
                        hipMemcpyDeviceToDevice));

  // Observations.  They are stored on the device in the order of their
  // first element, so that neighbouring lanes of the likelihood kernel read
  // neighbouring cells of a step's slab and neighbouring metadata; orig[k]
  // is the model's index of the k-th stored observation.
  {
    const int nobs = (int)m.obs.size();
    auto first_elem = [&](int i) -> uint64_t {
      const auto& e = m.obs[i].elem;
      return e.empty() ? ~0ull : ((uint64_t)(e[0] >> 31) << 32) | (e[0] & 0x7FFFFFFFu);
    };
    std::vector<int> ord(nobs), pos(nobs);
    for (int i = 0; i < nobs; i++) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return first_elem(a) < first_elem(b); });
    for (int k = 0; k < nobs; k++) pos[ord[k]] = k;
    std::vector<int> optr(1, 0), type, dtab, ospl;
    std::vector<uint32_t> elem;
    std::vector<double> y, sd, shape, factor, thresh, lc, iv;
    for (int k = 0; k < nobs; k++) {
      const auto& o = m.obs[ord[k]];
      elem.insert(elem.end(), o.elem.begin(), o.elem.end());
      for (uint32_t x : o.elem)
        if (x & 0x80000000u) c->obs_pop = true;
      optr.push_back((int)elem.size());
      type.push_back(o.type); dtab.push_back(o.datatable); ospl.push_back(o.ospl);
      y.push_back(o.y); sd.push_back(o.sd); shape.push_back(o.shape); factor.push_back(o.factor); thresh.push_back(o.thresh);
      // normal_probability (src/utils.cc:107-111) split into per-observation constants
      const double var = o.sd * o.sd;
      const bool normal = o.type == BEEP_OBS_NORMAL || o.type == BEEP_OBS_NORMAL_PERCENT;
      if (!normal || o.y == BEEP_OBS_UNKNOWN || o.y == BEEP_OBS_THRESH) c->obs_general = true;
      lc.push_back(normal ? -0.5 * std::log(2 * M_PI * var) : 0.0);
      iv.push_back(normal ? 1.0 / (2 * var) : 0.0);
    }
    beepdev::ObsDev& od = c->obs;
    od.nobs = nobs;
    od.ndt = m.ndatatable;
    od.TAG = (uint32_t)(m.T * m.AG);
    od.CAG = (uint32_t)(m.C * m.AG);
    if ((rc = upload(c, (int**)&od.ptr, optr)) || (rc = upload(c, (uint32_t**)&od.elem, elem)) ||
        (rc = upload(c, (int**)&od.type, type)) || (rc = upload(c, (int**)&od.dt, dtab)) ||
        (rc = upload(c, (double**)&od.y, y)) || (rc = upload(c, (double**)&od.sd, sd)) ||
        (rc = upload(c, (double**)&od.shape, shape)) || (rc = upload(c, (double**)&od.factor, factor)) ||
        (rc = upload(c, (double**)&od.thresh, thresh)) || (rc = upload(c, (int**)&od.orig, ord)) ||
        (rc = upload(c, (double**)&od.lc, lc)) || (rc = upload(c, (double**)&od.iv, iv)) ||
        (rc = upload(c, (int**)&od.ospl, ospl)))
      return rc;
    od.nospl = c->nospl;
    od.S = m.S;
    if ((rc = dalloc(c, (void**)&c->d_obs, sizeof(beepdev::ObsDev)))) return rc;
    HIPCHK(c, hipMemcpy(c->d_obs, &od, sizeof od, hipMemcpyHostToDevice));
    // Sections as lists of stored positions, ascending.
    std::vector<int> sptr(1, 0), sobs;
    for (const auto& l : m.section_obs) {
      std::vector<int> v;
      for (int i : l) v.push_back(pos[i]);
      std::sort(v.begin(), v.end());
      sobs.insert(sobs.end(), v.begin(), v.end());
      sptr.push_back((int)sobs.size());
    }
    c->sec_ptr = sptr;
    if ((rc = upload(c, &c->d_sec_ptr, sptr)) || (rc = upload(c, &c->d_sec_obs, sobs)) ||
        (rc = upload(c, &c->d_sec_ti, m.section_ti)) || (rc = upload(c, &c->d_sec_tf, m.section_tf)))
      return rc;
  }
  return BEEP_OK;
}

int beep_destroy(beep_ctx* c) {
  if (!c) return BEEP_EINVAL;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void* p : c->allocs) (void)hipFree(p);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
  for (auto& st : c->bstream)
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  if (c->bstart) (void)hipEventDestroy(c->bstart);
  for (auto& e : c->bend)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  if (c->h_stage64) (void)hipHostFree(c->h_stage64);
  delete c;
  return BEEP_OK;
}

int beep_path(const beep_ctx* c) {
  if (!c) return BEEP_EINVAL;
  if (c->stepwise) return c->spec ? BEEP_PATH_STEPWISE_SPEC : BEEP_PATH_STEPWISE;
  return c->spec ? BEEP_PATH_FUSED_SPEC : BEEP_PATH_FUSED;
}

int beep_get_dims(const beep_ctx* c, int32_t* d) {
  if (!c || !d) return BEEP_EINVAL;
  const auto& m = c->m;
  d[0] = m.A; d[1] = m.G; d[2] = m.C; d[3] = m.T; d[4] = m.S; d[5] = m.nparam;
  d[6] = (int)m.obs.size(); d[7] = m.ndatatable; d[8] = c->capacity; d[9] = c->nslot;
  return BEEP_OK;
}

size_t beep_device_bytes(const beep_ctx* c) { return c ? c->allocated : 0; }

}  // extern "C"

namespace {

// The FP32 slot block of derived tables (layout of TopoDev's offsets).
void fill_slot_block(const beep_ctx* c, const beephost::Tables& tb, float* block) {
  const TopoDev& tp = c->h_topo;
  std::fill(block, block + c->slot_stride, 0.f);
  for (size_t k = 0; k < tb.rate.size(); k++) block[tp.off_rate + k] = (float)tb.rate[k];
  for (size_t k = 0; k < tb.sus.size(); k++) block[tp.off_sus + k] = (float)tb.sus[k];
  for (size_t k = 0; k < tb.inf.size(); k++) block[tp.off_inf + k] = (float)tb.inf[k];
  for (size_t k = 0; k < tb.Ntime.size(); k++) block[tp.off_N + k] = (float)tb.Ntime[k];
  for (size_t k = 0; k < tb.efoi.size(); k++) block[tp.off_efoi + k] = (float)tb.efoi[k];
  for (size_t k = 0; k < tb.geo.size(); k++) block[tp.off_geo + k] = (float)tb.geo[k];
  for (size_t k = 0; k < tb.beta.size(); k++) block[tp.off_beta + k] = (float)tb.beta[k];
  if (tp.off_af >= 0)
    for (size_t k = 0; k < tb.af.size(); k++) block[tp.off_af + k] = (float)tb.af[k];
}

// Pinned host staging of at least nf floats and nd doubles.
int ensure_host_stage(beep_ctx* c, size_t nf, size_t nd) {
  if (nf > c->h_stage_cap) {
    if (c->h_stage) HIPCHK(c, hipHostFree(c->h_stage));
    c->h_stage = nullptr;
    c->h_stage_cap = 0;
    HIPCHK(c, hipHostMalloc((void**)&c->h_stage, nf * sizeof(float), hipHostMallocDefault));
    c->h_stage_cap = nf;
  }
  if (nd > c->h_stage64_cap) {
    if (c->h_stage64) HIPCHK(c, hipHostFree(c->h_stage64));
    c->h_stage64 = nullptr;
    c->h_stage64_cap = 0;
    HIPCHK(c, hipHostMalloc((void**)&c->h_stage64, nd * sizeof(double), hipHostMallocDefault));
    c->h_stage64_cap = nd;
  }
  return BEEP_OK;
}

// Runs f(i0, i1) over [0, n) on up to the host's hardware threads, in
// contiguous chunks of at least `grain` items (inline when one chunk).
template <class F>
void parallel_chunks(int n, int grain, F f) {
  int nt = (int)std::thread::hardware_concurrency();
  nt = std::max(1, std::min(nt, std::min(64, (n + grain - 1) / grain)));
  if (nt <= 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt);
  for (int t = 0; t < nt; t++) {
    const int i0 = (int)((long long)n * t / nt), i1 = (int)((long long)n * (t + 1) / nt);
    th.emplace_back([=, &f] { f(i0, i1); });
  }
  for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

// State::set_param (src/state.cc:66-91) for n parameter vectors into their
// slots.  The derivations (NGM, splines, area factors) run in parallel on the
// host threads into pinned staging; the slot blocks go to the device with one
// copy (slots contiguous and ascending) or one copy and a scatter kernel, and
// one synchronisation.  On any derivation error nothing is uploaded and the
// first failing vector's error is returned.
int beep_set_param(beep_ctx* c, int n, const int32_t* slot, const double* paramval, double* Pr_out,
                   int32_t* inb_out) {
  if (!c || n < 0 || (n > 0 && (!slot || !paramval))) return fail(c, BEEP_EINVAL, "beep_set_param: bad arguments");
  {
    std::vector<char> seen(c->nslot, 0);
    for (int i = 0; i < n; i++) {
      if (!valid_slot(c, slot[i])) return fail(c, BEEP_EINVAL, "slot %d out of range", slot[i]);
      if (seen[slot[i]]++) return fail(c, BEEP_EINVAL, "slot %d set twice in one call", slot[i]);
    }
  }
  if (n == 0) return BEEP_OK;
  const auto& m = c->m;
  const size_t stride = c->slot_stride, ostride = (size_t)c->nospl * m.S;
  int rc;
  if ((rc = ensure_host_stage(c, (size_t)n * stride, (size_t)n * ostride))) return rc;
  float* blk = c->h_stage;
  double* ob = c->h_stage64;
  std::vector<int> code(n, 0);
  std::vector<std::string> msg(n);
  parallel_chunks(n, 16, [&](int i0, int i1) {
    beephost::Tables tb;
    for (int i = i0; i < i1; i++) {
      const double* pv = paramval + (size_t)i * m.nparam;
      if (Pr_out) Pr_out[i] = m.prior_logpdf(pv);
      if (inb_out) inb_out[i] = m.inbounds(pv) ? 1 : 0;
      code[i] = m.derive(pv, tb, msg[i]);
      if (code[i]) continue;
      fill_slot_block(c, tb, blk + (size_t)i * stride);
      if (ostride) std::copy(tb.obs_spl.begin(), tb.obs_spl.end(), ob + (size_t)i * ostride);
    }
  });
  for (int i = 0; i < n; i++)
    if (code[i]) return fail(c, code[i], "set_param: %s", msg[i].c_str());
  bool contiguous = true;
  for (int i = 1; i < n && contiguous; i++) contiguous = slot[i] == slot[0] + i;
  if (contiguous) {
    HIPCHK(c, hipMemcpyAsync(c->d_tab + (size_t)slot[0] * stride, blk, (size_t)n * stride * sizeof(float),
                             hipMemcpyHostToDevice, c->stream));
    if (ostride)
      HIPCHK(c, hipMemcpyAsync(c->d_ospl + (size_t)slot[0] * ostride, ob, (size_t)n * ostride * sizeof(double),
                               hipMemcpyHostToDevice, c->stream));
  } else {
    const size_t fbytes = ((size_t)n * stride * sizeof(float) + 255) / 256 * 256;
    const size_t obytes = ((size_t)n * ostride * sizeof(double) + 255) / 256 * 256;
    if ((rc = ensure_scratch(c, fbytes + obytes + (size_t)n * sizeof(int)))) return rc;
    float* d_blk = (float*)c->d_scratch;
    double* d_ob = (double*)((char*)c->d_scratch + fbytes);
    int* d_slot = (int*)((char*)c->d_scratch + fbytes + obytes);
    HIPCHK(c, hipMemcpyAsync(d_blk, blk, (size_t)n * stride * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_slot, slot, (size_t)n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    const unsigned by = (unsigned)std::min<size_t>(64, (stride + 255) / 256);
    hipLaunchKernelGGL(beepdev::k_scatter_slots, dim3((unsigned)n, by), dim3(256), 0, c->stream, c->d_tab, stride,
                       d_blk, d_slot);
    HIPCHK(c, hipGetLastError());
    if (ostride) {
      HIPCHK(c, hipMemcpyAsync(d_ob, ob, (size_t)n * ostride * sizeof(double), hipMemcpyHostToDevice, c->stream));
      const unsigned oy = (unsigned)std::min<size_t>(64, (ostride + 255) / 256);
      hipLaunchKernelGGL(beepdev::k_scatter_rows64, dim3((unsigned)n, oy), dim3(256), 0, c->stream, c->d_ospl,
                         ostride, d_ob, d_slot);
      HIPCHK(c, hipGetLastError());
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));   // the staging is reused by the next call
  for (int i = 0; i < n; i++) {
    const double* pv = paramval + (size_t)i * m.nparam;
    std::copy(pv, pv + m.nparam, c->slot_param.begin() + (size_t)slot[i] * m.nparam);
    c->slot_set[slot[i]] = 1;
  }
  return BEEP_OK;
}

int beep_batch_set_param(int n, beep_ctx* const* ctx, const int32_t* slot, const double* const* paramval,
                         double* Pr_out, int32_t* inb_out) {
  if (n <= 0 || !ctx || !ctx[0] || !slot || !paramval) return BEEP_EINVAL;
  beep_ctx* c0 = ctx[0];
  std::vector<size_t> off(n + 1, 0);
  {
    std::map<beep_ctx*, std::vector<char>> seen;
    for (int i = 0; i < n; i++) {
      beep_ctx* c = ctx[i];
      if (!c || c->device != c0->device || !paramval[i]) return fail(c0, BEEP_EINVAL, "item %d: bad context or values", i);
      if (c->nospl) return fail(c0, BEEP_EINVAL, "item %d: a model with observation splines (use beep_set_param)", i);
      if (!valid_slot(c, slot[i])) return fail(c0, BEEP_EINVAL, "item %d: slot %d out of range", i, slot[i]);
      auto& sv = seen[c];
      if (sv.empty()) sv.assign(c->nslot, 0);
      if (sv[slot[i]]++) return fail(c0, BEEP_EINVAL, "item %d: slot %d set twice in one call", i, slot[i]);
      off[i + 1] = off[i] + c->slot_stride;
    }
  }
  int rc;
  const size_t total = off[n];
  if ((rc = ensure_host_stage(c0, total, 0))) return rc;
  float* blk = c0->h_stage;
  std::vector<int> code(n, 0);
  std::vector<std::string> msg(n);
  parallel_chunks(n, 16, [&](int i0, int i1) {
    beephost::Tables tb;
    for (int i = i0; i < i1; i++) {
      const auto& m = ctx[i]->m;
      const double* pv = paramval[i];
      if (Pr_out) Pr_out[i] = m.prior_logpdf(pv);
      if (inb_out) inb_out[i] = m.inbounds(pv) ? 1 : 0;
      code[i] = m.derive(pv, tb, msg[i]);
      if (!code[i]) fill_slot_block(ctx[i], tb, blk + off[i]);
    }
  });
  for (int i = 0; i < n; i++)
    if (code[i]) return fail(c0, code[i], "set_param (item %d): %s", i, msg[i].c_str());
  // Device staging: the packed blocks, then per item the source offset, the
  // destination slot and the length.
  const size_t pbytes = (total * sizeof(float) + 255) / 256 * 256;
  const size_t mbytes = (size_t)n * (8 + 8 + 4);
  if ((rc = ensure_scratch(c0, pbytes + mbytes))) return rc;
  std::vector<char> meta(mbytes);
  uint64_t* hs = (uint64_t*)meta.data();
  float** hd = (float**)(hs + n);
  uint32_t* hl = (uint32_t*)(hd + n);
  for (int i = 0; i < n; i++) {
    hs[i] = off[i];
    hd[i] = ctx[i]->d_tab + (size_t)slot[i] * ctx[i]->slot_stride;
    hl[i] = (uint32_t)ctx[i]->slot_stride;
  }
  char* d = (char*)c0->d_scratch;
  HIPCHK(c0, hipMemcpyAsync(d, blk, total * sizeof(float), hipMemcpyHostToDevice, c0->stream));
  HIPCHK(c0, hipMemcpyAsync(d + pbytes, meta.data(), mbytes, hipMemcpyHostToDevice, c0->stream));
  size_t maxlen = 0;
  for (int i = 0; i < n; i++) maxlen = std::max(maxlen, (size_t)ctx[i]->slot_stride);
  const unsigned by = (unsigned)std::min<size_t>(64, (maxlen + 255) / 256);
  hipLaunchKernelGGL(beepdev::k_scatter_blocks, dim3((unsigned)n, by), dim3(256), 0, c0->stream, (const float*)d,
                     (const uint64_t*)(d + pbytes), (float* const*)(d + pbytes + (size_t)8 * n),
                     (const uint32_t*)(d + pbytes + (size_t)16 * n));
  HIPCHK(c0, hipGetLastError());
  HIPCHK(c0, hipStreamSynchronize(c0->stream));   // staging reused by the next call
  for (int i = 0; i < n; i++) {
    beep_ctx* c = ctx[i];
    std::copy(paramval[i], paramval[i] + c->m.nparam, c->slot_param.begin() + (size_t)slot[i] * c->m.nparam);
    c->slot_set[slot[i]] = 1;
  }
  return BEEP_OK;
}

int beep_get_tables(beep_ctx* c, int slot, double* rate, double* sus, double* inf, double* Ntime, double* beta_af,
                    double* efoi, double* geo) {
  if (!c || !valid_slot(c, slot) || !c->slot_set[slot]) return fail(c, BEEP_EINVAL, "slot not set");
  beephost::Tables tb;
  std::string e;
  int rc = c->m.derive(&c->slot_param[(size_t)slot * c->m.nparam], tb, e);
  if (rc) return fail(c, rc, "%s", e.c_str());
  if (rate) std::copy(tb.rate.begin(), tb.rate.end(), rate);
  if (sus) std::copy(tb.sus.begin(), tb.sus.end(), sus);
  if (inf) std::copy(tb.inf.begin(), tb.inf.end(), inf);
  if (Ntime) std::copy(tb.Ntime.begin(), tb.Ntime.end(), Ntime);
  if (beta_af) {
    std::vector<double> ba = c->m.beta_af(tb);
    std::copy(ba.begin(), ba.end(), beta_af);
  }
  if (efoi) {
    std::vector<double> eg = c->m.efoi_cells(tb);
    std::copy(eg.begin(), eg.end(), efoi);
  }
  if (geo) std::copy(tb.geo.begin(), tb.geo.end(), geo);
  return BEEP_OK;
}

int beep_simulate(beep_ctx* c, int n, const int32_t* inst, const int32_t* slot, const uint32_t* key, int t0, int t1,
                  uint64_t seed, uint32_t stream, int flags) {
  if (!c) return BEEP_EINVAL;
  if (n < 0 || (n > 0 && (!inst || !slot || !key))) return fail(c, BEEP_EINVAL, "beep_simulate: bad arguments");
  if (t0 < 0 || t1 > c->m.S || t0 > t1) return fail(c, BEEP_EINVAL, "beep_simulate: bad step range [%d,%d)", t0, t1);
  if (stream > 0xFFFF) return fail(c, BEEP_EINVAL, "stream must be below 65536");
  int rc;
  if ((rc = check_instances(c, n, inst, nullptr, "beep_simulate"))) return rc;
  for (int i = 0; i < n; i++)
    if (!valid_slot(c, slot[i]) || !c->slot_set[slot[i]]) return fail(c, BEEP_EINVAL, "slot %d not set", slot[i]);
  if (n == 0 || t0 == t1) return BEEP_OK;
  if (flags & BEEP_SIM_PLAIN) {
    // The unbatched path: one launch per instance per step.
    std::vector<Job> jobs;
    jobs.reserve((size_t)n * (t1 - t0));
    for (int i = 0; i < n; i++)
      for (int s = t0; s < t1; s++) jobs.push_back(make_job(c, inst[i], slot[i], key[i], s, s + 1));
    if ((rc = put_jobs(c, jobs))) return rc;
    record(c, 0);
    for (size_t k = 0; k < jobs.size(); k++) {
      if (c->stepwise) rc = launch_stepwise(c, beepdev::MODE_SIM, c->d_jobs + k, 1, jobs[k].t0, jobs[k].t1, seed, stream);
      else rc = launch_fused(c, beepdev::MODE_SIM, c->d_jobs + k, 1, seed, stream);
      if (rc) return rc;
    }
    record(c, 1);
    c->timing[3] = (double)jobs.size();
  } else {
    // Longest jobs first: all instances of a context share one horizon, so
    // order is immaterial here; mixed batches sort by cost.
    std::vector<Job> jobs;
    jobs.reserve(n);
    for (int i = 0; i < n; i++) jobs.push_back(make_job(c, inst[i], slot[i], key[i], t0, t1));
    if ((rc = put_jobs(c, jobs))) return rc;
    record(c, 0);
    if (c->stepwise) rc = launch_stepwise(c, beepdev::MODE_SIM, c->d_jobs, n, t0, t1, seed, stream);
    else rc = launch_fused(c, beepdev::MODE_SIM, c->d_jobs, n, seed, stream);
    if (rc) return rc;
    record(c, 1);
    c->timing[3] = c->stepwise ? 2 * (t1 - t0) + 1 : 1;
  }
  rc = check_device_error(c);
  c->timing[0] = elapsed(c, 0, 1);
  return rc;
}

// beep_simulate's batched fused path with the job descriptors in `scratch`
// (this device's copy) instead of the context's job buffer: no device memory
// is allocated or freed.  The descriptors go up in chunks from a stack
// buffer, each copy waited for before the buffer is refilled; every
// descriptor the kernel reads is written by this call, so nothing needs
// clearing.  The call ends synchronised with its stream (the error word is
// read back), so calls from one host thread never overlap; calls that overlap
// on one device (other threads) race on `scratch`.
int beep_simulate_noalloc(beep_ctx* c, int n, const int32_t* inst, const int32_t* slot, const uint32_t* key, int t0,
                          int t1, uint64_t seed, uint32_t stream) {
  if (!c) return BEEP_EINVAL;
  if (n < 0 || (n > 0 && (!inst || !slot || !key))) return fail(c, BEEP_EINVAL, "beep_simulate_noalloc: bad arguments");
  if (t0 < 0 || t1 > c->m.S || t0 > t1)
    return fail(c, BEEP_EINVAL, "beep_simulate_noalloc: bad step range [%d,%d)", t0, t1);
  if (stream > 0xFFFF) return fail(c, BEEP_EINVAL, "stream must be below 65536");
  const long long max_jobs = BEEP_NOALLOC_BYTES / sizeof(Job);
  if (n > max_jobs)
    return fail(c, BEEP_EINVAL, "beep_simulate_noalloc: %d instances; at most %lld per call (%u bytes of job descriptors)",
                n, max_jobs, BEEP_NOALLOC_BYTES);
  if (c->stepwise)
    return fail(c, BEEP_EINVAL, "beep_simulate_noalloc: a step-wise context (its captured step loops read the "
                                "context's job buffer); use beep_simulate");
  int rc;
  if ((rc = check_instances(c, n, inst, nullptr, "beep_simulate_noalloc"))) return rc;
  for (int i = 0; i < n; i++)
    if (!valid_slot(c, slot[i]) || !c->slot_set[slot[i]]) return fail(c, BEEP_EINVAL, "slot %d not set", slot[i]);
  Job* d_jobs = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    d_jobs = (Job*)g_noalloc_ptr[c->device];
  }
  if (!d_jobs) return fail(c, BEEP_EHIP, "beep_simulate_noalloc: no device address of scratch on device %d", c->device);
  if (n == 0 || t0 == t1) return BEEP_OK;
  HIPCHK(c, hipSetDevice(c->device));
  constexpr int kChunk = 128;
  Job h[kChunk];
  for (int i0 = 0; i0 < n; i0 += kChunk) {
    const int m = std::min(kChunk, n - i0);
    for (int i = 0; i < m; i++) h[i] = make_job(c, inst[i0 + i], slot[i0 + i], key[i0 + i], t0, t1);
    HIPCHK(c, hipMemcpyAsync(d_jobs + i0, h, (size_t)m * sizeof(Job), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  record(c, 0);
  if ((rc = launch_fused(c, beepdev::MODE_SIM, d_jobs, n, seed, stream))) return rc;
  record(c, 1);
  c->timing[3] = 1;
  rc = check_device_error(c);
  c->timing[0] = elapsed(c, 0, 1);
  return rc;
}

int beep_mbp(beep_ctx* c, int n, const int32_t* init, const int32_t* prop, const int32_t* slot, const uint32_t* key,
             const uint8_t* simu_mask, uint64_t seed, uint32_t stream, int inf_update) {
  (void)inf_update;
  if (!c) return BEEP_EINVAL;
  if (!c->d_tn || !c->d_tm) return fail(c, BEEP_EINVAL, "MBPs need the context to keep transnum and transmean");
  if (n < 0 || (n > 0 && (!init || !prop || !slot || !key))) return fail(c, BEEP_EINVAL, "beep_mbp: bad arguments");
  if (stream > 0xFFFF) return fail(c, BEEP_EINVAL, "stream must be below 65536");
  int rc;
  // Proposals are distinct, and no proposal overwrites an initial state
  // another proposal of the call couples to.
  if ((rc = check_instances(c, n, prop, init, "beep_mbp"))) return rc;
  for (int i = 0; i < n; i++)
    if (!valid_slot(c, slot[i]) || !c->slot_set[slot[i]]) return fail(c, BEEP_EINVAL, "slot %d not set", slot[i]);
  if (n == 0) return BEEP_OK;
  const auto& m = c->m;
  uint8_t* d_mask = nullptr;
  size_t mask_sz = (size_t)m.S * m.A;
  if (simu_mask) {
    if ((rc = ensure_scratch(c, mask_sz * n))) return rc;
    d_mask = (uint8_t*)c->d_scratch;
    HIPCHK(c, hipMemcpyAsync(d_mask, simu_mask, mask_sz * n, hipMemcpyHostToDevice, c->stream));
  }
  std::vector<Job> jobs;
  jobs.reserve(n);
  for (int i = 0; i < n; i++) {
    Job j = make_job(c, prop[i], slot[i], key[i], 0, m.S);
    j.mode = beepdev::MODE_MBP;
    j.init_tn = c->d_tn + (size_t)init[i] * c->traj_sz;
    j.init_tm = c->d_tm + (size_t)init[i] * c->traj_sz;
    j.simu = d_mask ? d_mask + mask_sz * i : nullptr;
    jobs.push_back(j);
  }
  if ((rc = put_jobs(c, jobs))) return rc;
  record(c, 0);
  if (c->stepwise) rc = launch_stepwise(c, beepdev::MODE_MBP, c->d_jobs, n, 0, m.S, seed, stream);
  else rc = launch_fused(c, beepdev::MODE_MBP, c->d_jobs, n, seed, stream);
  if (rc) return rc;
  record(c, 1);
  rc = check_device_error(c);
  c->timing[0] = elapsed(c, 0, 1);
  c->timing[3] = 1;
  return rc;
}

int beep_obs_loglik(beep_ctx* c, int n, const int32_t* inst, int section, double* L_out, double* EF_dt_out) {
  if (!c) return BEEP_EINVAL;
  if (n < 0 || (n > 0 && (!inst || !L_out))) return fail(c, BEEP_EINVAL, "beep_obs_loglik: bad arguments");
  if (!c->d_tn) return fail(c, BEEP_EINVAL, "the observation model needs transnum to be kept");
  if (section >= (int)c->m.section_obs.size()) return fail(c, BEEP_EINVAL, "section %d out of range", section);
  for (int i = 0; i < n; i++)
    if (!valid_inst(c, inst[i])) return fail(c, BEEP_EINVAL, "instance %d out of range", inst[i]);
  if (n == 0) return BEEP_OK;
  int rc;
  if (c->nospl) {
    for (int i = 0; i < n; i++)
      if (c->inst_slot[inst[i]] < 0)
        return fail(c, BEEP_EINVAL, "instance %d has no parameter slot for its observation splines", inst[i]);
    if ((rc = sync_inst_slot(c))) return rc;
  }
  if (section < 0 && !EF_dt_out) {
    // The whole observation model: the mixed batches' chunked kernels, so a
    // batch and this call give the same sums.
    std::vector<beep_ctx*> cs(1, c);
    std::vector<const int32_t*> ip(1, inst);
    std::vector<int> ni(1, n);
    if ((rc = obs_loglik_chunked(c, cs, ip, ni, true, L_out, 2))) return rc;
    c->timing[1] = elapsed(c, 2, 3);
    for (int i = 0; i < n; i++)
      if (std::isnan(L_out[i])) return fail(c, BEEP_ENAN, "NaN observation log-likelihood");   // obsmodel.cc:32
    return BEEP_OK;
  }
  const int ndt = c->m.ndatatable;
  size_t out_doubles = (size_t)n * (1 + ndt);
  if ((rc = ensure_scratch(c, out_doubles * sizeof(double) + (size_t)n * sizeof(int)))) return rc;
  double* d_out = (double*)c->d_scratch;
  int* d_inst = (int*)(d_out + out_doubles);
  HIPCHK(c, hipMemcpyAsync(d_inst, inst, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  const int* sobs = nullptr;
  int nsel = c->m.obs.size();
  if (section >= 0) {
    sobs = c->d_sec_obs + c->sec_ptr[section];
    nsel = c->sec_ptr[section + 1] - c->sec_ptr[section];
  }
  record(c, 2);
  launch_obs_loglik(c, n, d_inst, sobs, nsel, d_out, section < 0 ? 1 : 0, EF_dt_out != nullptr);
  HIPCHK(c, hipGetLastError());
  record(c, 3);
  std::vector<double> h(out_doubles);
  HIPCHK(c, hipMemcpyAsync(h.data(), d_out, out_doubles * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->timing[1] = elapsed(c, 2, 3);
  for (int i = 0; i < n; i++) {
    L_out[i] = h[(size_t)i * (1 + ndt)];
    if (std::isnan(L_out[i])) return fail(c, BEEP_ENAN, "NaN observation log-likelihood");   // obsmodel.cc:32
    if (EF_dt_out)
      for (int d = 0; d < ndt; d++) EF_dt_out[(size_t)i * ndt + d] = h[(size_t)i * (1 + ndt) + 1 + d];
  }
  return BEEP_OK;
}

int beep_obs_values(beep_ctx* c, int inst, double* value_out) {
  if (!c || !valid_inst(c, inst) || !value_out) return fail(c, BEEP_EINVAL, "beep_obs_values: bad arguments");
  if (!c->d_tn) return fail(c, BEEP_EINVAL, "the observation model needs transnum to be kept");
  int nobs = c->m.obs.size();
  if (nobs == 0) return BEEP_OK;
  int rc;
  if ((rc = ensure_scratch(c, (size_t)nobs * sizeof(double)))) return rc;
  double* d_out = (double*)c->d_scratch;
  if (c->nospl && c->inst_slot[inst] < 0)
    return fail(c, BEEP_EINVAL, "instance %d has no parameter slot for its observation splines", inst);
  const double* splv = c->nospl ? c->d_ospl + (size_t)c->inst_slot[inst] * c->nospl * c->m.S : nullptr;
  hipLaunchKernelGGL(beepdev::k_obs_values, dim3((nobs + 255) / 256), dim3(256), 0, c->stream, c->d_obs,
                     c->d_tn + (size_t)inst * c->traj_sz, c->d_pt ? c->d_pt + (size_t)inst * c->pop_traj_sz : nullptr,
                     splv, d_out);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(value_out, d_out, nobs * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BEEP_OK;
}

int beep_pois_loglik(beep_ctx* c, int n, const int32_t* inst, double* L_out) {
  if (!c) return BEEP_EINVAL;
  if (n < 0 || (n > 0 && (!inst || !L_out))) return fail(c, BEEP_EINVAL, "beep_pois_loglik: bad arguments");
  if (!c->d_tn || !c->d_tm) return fail(c, BEEP_EINVAL, "State::likelihood needs transnum and transmean");
  for (int i = 0; i < n; i++)
    if (!valid_inst(c, inst[i])) return fail(c, BEEP_EINVAL, "instance %d out of range", inst[i]);
  if (n == 0) return BEEP_OK;
  int rc;
  if ((rc = ensure_scratch(c, (size_t)n * (sizeof(double) + sizeof(int))))) return rc;
  double* d_out = (double*)c->d_scratch;
  int* d_inst = (int*)(d_out + n);
  HIPCHK(c, hipMemcpyAsync(d_inst, inst, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  record(c, 4);
  hipLaunchKernelGGL(beepdev::k_pois_loglik, dim3(n), dim3(256), 0, c->stream, d_inst, c->d_tn, c->d_tm, c->traj_sz,
                     d_out);
  HIPCHK(c, hipGetLastError());
  record(c, 5);
  HIPCHK(c, hipMemcpyAsync(L_out, d_out, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->timing[2] = elapsed(c, 4, 5);
  return BEEP_OK;
}

int beep_accept(beep_ctx* c, int n, const int32_t* init, const int32_t* prop, int mode, double invT, double EFcut,
                const double* dlogq, const uint32_t* key, uint64_t seed, uint32_t stream, double* al_out,
                uint8_t* accepted) {
  if (!c) return BEEP_EINVAL;
  if (n < 0 || (n > 0 && (!init || !prop || !dlogq || !key || !al_out || !accepted)) || (mode != 0 && mode != 1))
    return fail(c, BEEP_EINVAL, "beep_accept: bad arguments");
  for (int i = 0; i < n; i++)
    if (!valid_inst(c, init[i]) || !valid_inst(c, prop[i])) return fail(c, BEEP_EINVAL, "instance out of range");
  if (n == 0) return BEEP_OK;
  int rc;
  size_t bytes = (size_t)n * (2 * sizeof(int) + 2 * sizeof(double) + sizeof(uint32_t) + sizeof(double) + 1);
  if ((rc = ensure_scratch(c, bytes + 64))) return rc;
  char* p = (char*)c->d_scratch;
  double* d_dlogq = (double*)p; p += n * sizeof(double);
  double* d_al = (double*)p; p += n * sizeof(double);
  int* d_init = (int*)p; p += n * sizeof(int);
  int* d_prop = (int*)p; p += n * sizeof(int);
  uint32_t* d_key = (uint32_t*)p; p += n * sizeof(uint32_t);
  uint8_t* d_acc = (uint8_t*)p;
  HIPCHK(c, hipMemcpyAsync(d_dlogq, dlogq, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_init, init, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_prop, prop, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_key, key, n * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(beepdev::k_accept, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, d_init, d_prop, c->d_EF,
                     d_dlogq, d_key, mode, invT, EFcut, seed, stream, d_al, d_acc);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(al_out, d_al, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(accepted, d_acc, n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; i++)
    if (std::isnan(al_out[i])) return fail(c, BEEP_ENAN, "NaN acceptance probability");   // src/mbp.cc:353
  return BEEP_OK;
}

// ---- state transfer ----

int beep_get_transnum(beep_ctx* c, int inst, int32_t* out) {
  if (!c || !valid_inst(c, inst) || !out || !c->d_tn) return fail(c, BEEP_EINVAL, "beep_get_transnum: bad arguments");
  const auto& m = c->m;
  std::vector<int> h(c->traj_sz);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->d_tn + (size_t)inst * c->traj_sz, c->traj_sz * sizeof(int),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int s = 0; s < m.S; s++)
    for (int a = 0; a < m.A; a++)
      for (int t = 0; t < m.T; t++)
        for (int g = 0; g < m.G; g++)
          out[(((size_t)s * m.A + a) * m.T + t) * m.G + g] = h[((size_t)s * m.T + t) * m.AG + a * m.G + g];
  return BEEP_OK;
}

int beep_get_transmean(beep_ctx* c, int inst, float* out) {
  if (!c || !valid_inst(c, inst) || !out || !c->d_tm) return fail(c, BEEP_EINVAL, "beep_get_transmean: bad arguments");
  const auto& m = c->m;
  std::vector<float> h(c->traj_sz);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->d_tm + (size_t)inst * c->traj_sz, c->traj_sz * sizeof(float),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int s = 0; s < m.S; s++)
    for (int a = 0; a < m.A; a++)
      for (int t = 0; t < m.T; t++)
        for (int g = 0; g < m.G; g++)
          out[(((size_t)s * m.A + a) * m.T + t) * m.G + g] = h[((size_t)s * m.T + t) * m.AG + a * m.G + g];
  return BEEP_OK;
}

int beep_get_pop(beep_ctx* c, int inst, int32_t* out) {
  if (!c || !valid_inst(c, inst) || !out || !c->d_pt) return fail(c, BEEP_EINVAL, "beep_get_pop: bad arguments");
  const auto& m = c->m;
  std::vector<int> h(c->pop_traj_sz);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->d_pt + (size_t)inst * c->pop_traj_sz, c->pop_traj_sz * sizeof(int),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int s = 0; s < m.S; s++)
    for (int a = 0; a < m.A; a++)
      for (int co = 0; co < m.C; co++)
        for (int g = 0; g < m.G; g++)
          out[(((size_t)s * m.A + a) * m.C + co) * m.G + g] = h[((size_t)s * m.C + co) * m.AG + a * m.G + g];
  return BEEP_OK;
}

int beep_get_state(beep_ctx* c, int inst, int32_t* out) {
  if (!c || !valid_inst(c, inst) || !out) return fail(c, BEEP_EINVAL, "beep_get_state: bad arguments");
  const auto& m = c->m;
  std::vector<int> h(c->state_sz);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->d_state + (size_t)inst * c->state_sz, c->state_sz * sizeof(int),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int a = 0; a < m.A; a++)
    for (int co = 0; co < m.C; co++)
      for (int g = 0; g < m.G; g++) out[((size_t)a * m.C + co) * m.G + g] = h[(size_t)co * m.AG + a * m.G + g];
  return BEEP_OK;
}

int beep_put_state(beep_ctx* c, int inst, const int32_t* pop) {
  if (!c || !valid_inst(c, inst) || !pop) return fail(c, BEEP_EINVAL, "beep_put_state: bad arguments");
  const auto& m = c->m;
  std::vector<int> h(c->state_sz);
  for (int a = 0; a < m.A; a++)
    for (int co = 0; co < m.C; co++)
      for (int g = 0; g < m.G; g++) h[(size_t)co * m.AG + a * m.G + g] = pop[((size_t)a * m.C + co) * m.G + g];
  HIPCHK(c, hipMemcpyAsync(c->d_state + (size_t)inst * c->state_sz, h.data(), c->state_sz * sizeof(int),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BEEP_OK;
}

int beep_put_transnum(beep_ctx* c, int inst, int slot, const int32_t* tn) {
  if (!c || !valid_inst(c, inst) || !tn || !c->d_tn) return fail(c, BEEP_EINVAL, "beep_put_transnum: bad arguments");
  if (!valid_slot(c, slot) || !c->slot_set[slot]) return fail(c, BEEP_EINVAL, "slot %d not set", slot);
  const auto& m = c->m;
  std::vector<int> h(c->traj_sz);
  for (int s = 0; s < m.S; s++)
    for (int a = 0; a < m.A; a++)
      for (int t = 0; t < m.T; t++)
        for (int g = 0; g < m.G; g++)
          h[((size_t)s * m.T + t) * m.AG + a * m.G + g] = tn[(((size_t)s * m.A + a) * m.T + t) * m.G + g];
  HIPCHK(c, hipMemcpyAsync(c->d_tn + (size_t)inst * c->traj_sz, h.data(), c->traj_sz * sizeof(int),
                           hipMemcpyHostToDevice, c->stream));
  std::vector<Job> jobs(1, make_job(c, inst, slot, 0, 0, m.S));
  jobs[0].mode = beepdev::MODE_REPLAY;
  int rc;
  if ((rc = put_jobs(c, jobs))) return rc;
  if (c->stepwise) rc = launch_stepwise(c, beepdev::MODE_REPLAY, c->d_jobs, 1, 0, m.S, 0, 0);
  else rc = launch_fused(c, beepdev::MODE_REPLAY, c->d_jobs, 1, 0, 0);
  if (rc) return rc;
  return check_device_error(c);
}

int beep_copy_instances(beep_ctx* c, int n, const int32_t* dst, const int32_t* src) {
  if (!c || n < 0 || (n > 0 && (!dst || !src))) return fail(c, BEEP_EINVAL, "beep_copy_instances: bad arguments");
  for (int i = 0; i < n; i++) {
    if (!valid_inst(c, dst[i]) || !valid_inst(c, src[i])) return fail(c, BEEP_EINVAL, "instance out of range");
    if (dst[i] == src[i]) continue;
    HIPCHK(c, hipMemcpyAsync(c->d_state + (size_t)dst[i] * c->state_sz, c->d_state + (size_t)src[i] * c->state_sz,
                             c->state_sz * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
    if (c->d_tn)
      HIPCHK(c, hipMemcpyAsync(c->d_tn + (size_t)dst[i] * c->traj_sz, c->d_tn + (size_t)src[i] * c->traj_sz,
                               c->traj_sz * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
    if (c->d_tm)
      HIPCHK(c, hipMemcpyAsync(c->d_tm + (size_t)dst[i] * c->traj_sz, c->d_tm + (size_t)src[i] * c->traj_sz,
                               c->traj_sz * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    if (c->d_pt)
      HIPCHK(c, hipMemcpyAsync(c->d_pt + (size_t)dst[i] * c->pop_traj_sz, c->d_pt + (size_t)src[i] * c->pop_traj_sz,
                               c->pop_traj_sz * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_EF + dst[i], c->d_EF + src[i], sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    c->inst_slot[dst[i]] = c->inst_slot[src[i]];     // the trajectory's parameters go with it
    c->inst_slot_dirty = true;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BEEP_OK;
}

int beep_copy_states(beep_ctx* c, int n, const int32_t* dst, const int32_t* src) {
  if (!c || n < 0 || (n > 0 && (!dst || !src))) return fail(c, BEEP_EINVAL, "beep_copy_states: bad arguments");
  for (int i = 0; i < n; i++) {
    if (!valid_inst(c, dst[i]) || !valid_inst(c, src[i])) return fail(c, BEEP_EINVAL, "instance out of range");
    if (dst[i] == src[i]) continue;
    HIPCHK(c, hipMemcpyAsync(c->d_state + (size_t)dst[i] * c->state_sz, c->d_state + (size_t)src[i] * c->state_sz,
                             c->state_sz * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BEEP_OK;
}

int beep_derive_slots(beep_ctx* c, int n, const int32_t* dst, const int32_t* src, const double* factor,
                      const int32_t* from_step) {
  if (!c || n < 0 || (n > 0 && (!dst || !src || !factor || !from_step)))
    return fail(c, BEEP_EINVAL, "beep_derive_slots: bad arguments");
  if (n == 0) return BEEP_OK;
  std::vector<float> f(n);
  for (int i = 0; i < n; i++) {
    if (!valid_slot(c, dst[i]) || !valid_slot(c, src[i]) || !c->slot_set[src[i]])
      return fail(c, BEEP_EINVAL, "slot out of range or not set");
    if (from_step[i] < 0 || from_step[i] > c->m.S) return fail(c, BEEP_EINVAL, "from_step out of range");
    if (!(factor[i] >= 0) || !std::isfinite(factor[i])) return fail(c, BEEP_EINVAL, "bad factor");
    f[i] = (float)factor[i];
  }
  int rc;
  size_t bytes = (size_t)n * (3 * sizeof(int) + sizeof(float));
  if ((rc = ensure_scratch(c, bytes))) return rc;
  int* d_dst = (int*)c->d_scratch;
  int* d_src = d_dst + n;
  int* d_from = d_src + n;
  float* d_f = (float*)(d_from + n);
  HIPCHK(c, hipMemcpyAsync(d_dst, dst, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_src, src, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_from, from_step, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_f, f.data(), n * sizeof(float), hipMemcpyHostToDevice, c->stream));
  int bx = (int)std::min<size_t>(64, (c->slot_stride + 255) / 256);
  hipLaunchKernelGGL(beepdev::k_derive_slot, dim3(n, bx), dim3(256), 0, c->stream, c->d_tab, c->slot_stride, d_dst,
                     d_src, d_f, d_from, c->h_topo.off_beta, c->m.S, c->m.nR);
  HIPCHK(c, hipGetLastError());
  const size_t osz = (size_t)c->nospl * c->m.S;     // observation splines: the source slot's
  for (int i = 0; c->nospl && i < n; i++)
    HIPCHK(c, hipMemcpyAsync(c->d_ospl + (size_t)dst[i] * osz, c->d_ospl + (size_t)src[i] * osz, osz * sizeof(double),
                             hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; i++) {
    c->slot_set[dst[i]] = 1;
    std::copy(c->slot_param.begin() + (size_t)src[i] * c->m.nparam,
              c->slot_param.begin() + (size_t)(src[i] + 1) * c->m.nparam,
              c->slot_param.begin() + (size_t)dst[i] * c->m.nparam);
  }
  return BEEP_OK;
}

int beep_sum_transnum(beep_ctx* c, int n, const int32_t* inst, int tr, int t0, int t1, double* out) {
  if (!c || n < 0 || (n > 0 && (!inst || !out)) || !c->d_tn) return fail(c, BEEP_EINVAL, "beep_sum_transnum: bad arguments");
  if (tr < 0 || tr >= c->m.T || t0 < 0 || t1 > c->m.S || t0 >= t1) return fail(c, BEEP_EINVAL, "bad transition or steps");
  for (int i = 0; i < n; i++)
    if (!valid_inst(c, inst[i])) return fail(c, BEEP_EINVAL, "instance out of range");
  if (n == 0) return BEEP_OK;
  int rc;
  size_t nst = (size_t)(t1 - t0);
  if ((rc = ensure_scratch(c, (size_t)n * nst * sizeof(double) + (size_t)n * sizeof(int)))) return rc;
  double* d_out = (double*)c->d_scratch;
  int* d_inst = (int*)(d_out + (size_t)n * nst);
  HIPCHK(c, hipMemcpyAsync(d_inst, inst, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(beepdev::k_sum_transnum, dim3((unsigned)nst, n), dim3(256), 0, c->stream, d_inst, c->d_tn,
                     c->traj_sz, c->m.AG, c->m.T, tr, t0, (int)nst, d_out);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(out, d_out, (size_t)n * nst * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BEEP_OK;
}

int beep_timing(beep_ctx* c, double* out) {
  if (!c || !out) return BEEP_EINVAL;
  for (int k = 0; k < 4; k++) out[k] = c->timing[k];
  return BEEP_OK;
}

// ---- PMCMC particle filter ----

int beep_pmcmc_obs_prob(beep_ctx* c, int n, const int32_t* inst, int slot, const uint32_t* key, double invT,
                        uint64_t seed, uint32_t stream, double* obprob_out, int32_t* backpart_out) {
  if (!c) return BEEP_EINVAL;
  if (n <= 0 || !inst || !key || !obprob_out) return fail(c, BEEP_EINVAL, "beep_pmcmc_obs_prob: bad arguments");
  if (!c->d_tn) return fail(c, BEEP_EINVAL, "the observation model needs transnum to be kept");
  if (stream > 0xFFFF) return fail(c, BEEP_EINVAL, "stream must be below 65536");
  if (!valid_slot(c, slot) || !c->slot_set[slot]) return fail(c, BEEP_EINVAL, "slot %d not set", slot);
  if (!(invT >= 0) || !std::isfinite(invT)) return fail(c, BEEP_EINVAL, "bad invT");
  const auto& m = c->m;
  const int nsec = (int)m.section_ti.size();
  if (nsec == 0) return fail(c, BEEP_EINVAL, "the model has no observation sections");
  for (int k = 0; k < nsec; k++)
    if (m.section_ti[k] != (k ? m.section_tf[k - 1] : 0) || (k == nsec - 1 && m.section_tf[k] != m.S))
      return fail(c, BEEP_EINVAL, "the observation sections must tile [0, S) in order");
  std::vector<char> seen(c->capacity, 0);
  for (int i = 0; i < n; i++) {
    if (!valid_inst(c, inst[i])) return fail(c, BEEP_EINVAL, "instance %d out of range", inst[i]);
    if (seen[inst[i]]++) return fail(c, BEEP_EINVAL, "instance %d appears twice", inst[i]);
  }
  int rc;
  // Persistent ancestry + instance list (read later by particle_sample).
  size_t pf_ints = (size_t)(nsec + 1) * n + n;
  if (pf_ints > c->pf_cap) {
    if (c->d_pf) {
      HIPCHK(c, hipFree(c->d_pf));
      c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), (void*)c->d_pf));
      c->allocated -= c->pf_cap * sizeof(int);
      c->d_pf = nullptr;
      c->pf_cap = 0;
    }
    if ((rc = dalloc(c, (void**)&c->d_pf, pf_ints * sizeof(int)))) return rc;
    c->pf_cap = pf_ints;
  }
  c->pf_n = 0;
  int* d_bp = c->d_pf;
  int* d_inst = c->d_pf + (size_t)(nsec + 1) * n;
  std::vector<int> head((size_t)n);
  for (int i = 0; i < n; i++) head[i] = i;
  HIPCHK(c, hipMemcpyAsync(d_bp, head.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_inst, inst, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  // Scratch: L rows, bootstrap work arrays, the estimate.
  const int Lstride = 1 + m.ndatatable;
  size_t bytes = (size_t)n * Lstride * sizeof(double) + (size_t)n * sizeof(double) + 3 * (size_t)n * sizeof(int) + 64;
  if ((rc = ensure_scratch(c, bytes))) return rc;
  double* d_ob = (double*)c->d_scratch;
  double* d_L = d_ob + 1;
  double* d_w = d_L + (size_t)n * Lstride;
  int* d_first = (int*)(d_w + n);
  int* d_draw = d_first + n;
  int* d_extra = d_draw + n;
  HIPCHK(c, hipMemsetAsync(d_ob, 0, sizeof(double), c->stream));
  // Jobs of every section.
  std::vector<Job> jobs;
  jobs.reserve((size_t)nsec * n);
  for (int k = 0; k < nsec; k++)
    for (int i = 0; i < n; i++) jobs.push_back(make_job(c, inst[i], slot, key[i], m.section_ti[k], m.section_tf[k]));
  if ((rc = put_jobs(c, jobs))) return rc;
  if ((rc = sync_inst_slot(c))) return rc;
  const unsigned mv_x = (unsigned)std::min<size_t>(16, (c->state_sz + 255) / 256);
  record(c, 0);
  for (int k = 0; k < nsec; k++) {
    if (k > 0) {
      hipLaunchKernelGGL(beepdev::k_move_states, dim3(n, mv_x), dim3(256), 0, c->stream, d_inst, d_bp + (size_t)k * n,
                         c->d_state, c->state_sz);
      HIPCHK(c, hipGetLastError());
    }
    const Job* dj = c->d_jobs + (size_t)k * n;
    if (c->stepwise) rc = launch_stepwise(c, beepdev::MODE_SIM, dj, n, m.section_ti[k], m.section_tf[k], seed, stream);
    else rc = launch_fused(c, beepdev::MODE_SIM, dj, n, seed, stream);
    if (rc) return rc;
    const int nsel = c->sec_ptr[k + 1] - c->sec_ptr[k];
    launch_obs_loglik(c, n, d_inst, c->d_sec_obs + c->sec_ptr[k], nsel, d_L, 0, false);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(beepdev::k_bootstrap, dim3(1), dim3(BEEP_BOOT_BLOCK), 0, c->stream, n, d_L, Lstride, invT, seed,
                       stream, k, d_bp + (size_t)(k + 1) * n, d_w, d_first, d_draw, d_extra, d_ob, c->d_err);
    HIPCHK(c, hipGetLastError());
  }
  record(c, 1);
  double ob = 0;
  HIPCHK(c, hipMemcpyAsync(&ob, d_ob, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (backpart_out)
    HIPCHK(c, hipMemcpyAsync(backpart_out, d_bp, (size_t)(nsec + 1) * n * sizeof(int), hipMemcpyDeviceToHost,
                             c->stream));
  rc = check_device_error(c);      // synchronises the stream
  c->timing[0] = elapsed(c, 0, 1);
  c->timing[3] = 3.0 * nsec - 1 + (c->stepwise ? 2.0 * m.S + nsec : nsec);   // kernel launches
  if (rc) return rc;
  *obprob_out = ob;
  c->pf_n = n;
  c->pf_slot = slot;
  return BEEP_OK;
}

int beep_pmcmc_particle_sample(beep_ctx* c, int dst) {
  if (!c) return BEEP_EINVAL;
  if (!valid_inst(c, dst)) return fail(c, BEEP_EINVAL, "instance %d out of range", dst);
  if (c->pf_n <= 0) return fail(c, BEEP_EINVAL, "no particle filter run to sample from");
  const auto& m = c->m;
  const int nsec = (int)m.section_ti.size(), n = c->pf_n;
  const int* d_inst = c->d_pf + (size_t)(nsec + 1) * n;
  size_t maxlen = 0;
  for (int k = 0; k < nsec; k++) maxlen = std::max<size_t>(maxlen, (size_t)(m.section_tf[k] - m.section_ti[k]));
  const unsigned gx = (unsigned)std::min<size_t>(64, (maxlen * m.T * m.AG + 255) / 256);
  hipLaunchKernelGGL(beepdev::k_particle_sample, dim3(gx, nsec), dim3(256), 0, c->stream, n, nsec, c->d_sec_ti,
                     c->d_sec_tf, c->d_pf, d_inst, dst, c->d_tn, c->d_tm, c->d_pt, c->traj_sz, c->pop_traj_sz,
                     m.T * m.AG, m.C * m.AG);
  HIPCHK(c, hipGetLastError());
  // EF of the sampled particle is not defined by the section run.
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->inst_slot[dst] = c->pf_slot;                   // the filter's parameters
  c->inst_slot_dirty = true;
  return BEEP_OK;
}

int beep_abcsmc_weights(int device, int nvar, int n_prev, const double* prev, const double* w_prev, int n_new,
                        const double* theta, const double* inv_M, double size, const double* logprior,
                        double* w_out) {
  if (nvar <= 0 || nvar > 32 || n_prev <= 0 || n_new < 0 || !prev || !w_prev || !inv_M || !(size > 0) ||
      (n_new > 0 && (!theta || !logprior || !w_out)))
    return BEEP_EINVAL;
  if (n_new == 0) return BEEP_OK;
  if (hipSetDevice(device) != hipSuccess) return BEEP_EHIP;
  const size_t b_prev = (size_t)n_prev * nvar * 8, b_w = (size_t)n_prev * 8, b_th = (size_t)n_new * nvar * 8,
               b_M = (size_t)nvar * nvar * 8, b_lp = (size_t)n_new * 8, b_out = (size_t)n_new * 8;
  char* d = nullptr;
  if (hipMalloc((void**)&d, b_prev + b_w + b_th + b_M + b_lp + b_out) != hipSuccess) return BEEP_ENOMEM;
  double* d_prev = (double*)d;
  double* d_w = (double*)(d + b_prev);
  double* d_th = (double*)(d + b_prev + b_w);
  double* d_M = (double*)(d + b_prev + b_w + b_th);
  double* d_lp = (double*)(d + b_prev + b_w + b_th + b_M);
  double* d_out = (double*)(d + b_prev + b_w + b_th + b_M + b_lp);
  int rc = BEEP_OK;
  if (hipMemcpy(d_prev, prev, b_prev, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_w, w_prev, b_w, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_th, theta, b_th, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_M, inv_M, b_M, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_lp, logprior, b_lp, hipMemcpyHostToDevice) != hipSuccess)
    rc = BEEP_EHIP;
  if (!rc) {
    const size_t lds = 8 * ((size_t)nvar * nvar + (size_t)nvar * BEEP_ABC_BLOCK + (size_t)BEEP_ABC_TILE * nvar +
                            BEEP_ABC_TILE);
    if (lds > 65536 &&
        hipFuncSetAttribute((const void*)beepdev::k_abcsmc_weights, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds) != hipSuccess)
      rc = BEEP_EHIP;
    if (!rc) {
      hipLaunchKernelGGL(beepdev::k_abcsmc_weights, dim3((n_new + BEEP_ABC_BLOCK - 1) / BEEP_ABC_BLOCK),
                         dim3(BEEP_ABC_BLOCK), lds, 0, nvar, n_prev, d_prev, d_w, n_new, d_th, d_M, size, d_lp, d_out);
      if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = BEEP_EHIP;
    }
  }
  if (!rc && hipMemcpy(w_out, d_out, b_out, hipMemcpyDeviceToHost) != hipSuccess) rc = BEEP_EHIP;
  (void)hipFree(d);
  return rc;
}

int beep_mc3_swap(beep_ctx* c, int n, const int32_t* inst, const int32_t* run, const double* invT, uint64_t seed,
                  uint32_t stream, int32_t* order_out, int32_t* ntr_out, int32_t* nac_out) {
  if (!c) return BEEP_EINVAL;
  if (n <= 0 || !inst || !run || !invT || !order_out || !ntr_out || !nac_out || stream > 0xFFFF)
    return fail(c, BEEP_EINVAL, "beep_mc3_swap: bad arguments");
  for (int i = 0; i < n; i++)
    if (!valid_inst(c, inst[i])) return fail(c, BEEP_EINVAL, "instance %d out of range", inst[i]);
  int rc;
  // scratch: invT[n], EF[n] (double); inst[n], run[n], order[n], ntr[n], nac[n] (int)
  if ((rc = ensure_scratch(c, (size_t)n * (2 * 8 + 5 * 4) + 64))) return rc;
  double* d_invT = (double*)c->d_scratch;
  double* d_EF = d_invT + n;
  int* d_inst = (int*)(d_EF + n);
  int* d_run = d_inst + n;
  int* d_order = d_run + n;
  int* d_ntr = d_order + n;
  int* d_nac = d_ntr + n;
  HIPCHK(c, hipMemcpyAsync(d_invT, invT, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_inst, inst, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_run, run, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(beepdev::k_mc3_swap, dim3(1), dim3(64), 0, c->stream, n, d_run, d_invT, c->d_EF, d_inst, d_EF,
                     d_order, d_ntr, d_nac, seed, stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(order_out, d_order, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(ntr_out, d_ntr, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(nac_out, d_nac, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BEEP_OK;
}

int beep_bootstrap(int device, int n, const double* L, double invT, uint64_t seed, uint32_t stream, int sec,
                   int32_t* bp_out, double* obprob_out) {
  if (n <= 0 || !L || !bp_out || !obprob_out || stream > 0xFFFF || sec < 0 || !(invT >= 0) || !std::isfinite(invT))
    return BEEP_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return BEEP_EHIP;
  const size_t bytes = 16 + (size_t)n * (8 + 8 + 4 * 4);     // estimate, error flag, L, wsum, 4 int arrays
  char* d = nullptr;
  if (hipMalloc((void**)&d, bytes) != hipSuccess) return BEEP_ENOMEM;
  double* d_ob = (double*)d;
  int* d_err = (int*)(d + 8);
  double* d_L = (double*)(d + 16);
  double* d_w = d_L + n;
  int* d_bp = (int*)(d_w + n);
  int* d_first = d_bp + n;
  int* d_draw = d_first + n;
  int* d_extra = d_draw + n;
  int rc = BEEP_OK, err = 0;
  if (hipMemset(d, 0, 16) != hipSuccess || hipMemcpy(d_L, L, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    rc = BEEP_EHIP;
  if (!rc) {
    hipLaunchKernelGGL(beepdev::k_bootstrap, dim3(1), dim3(BEEP_BOOT_BLOCK), 0, 0, n, d_L, 1, invT, seed, stream, sec,
                       d_bp, d_w, d_first, d_draw, d_extra, d_ob, d_err);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = BEEP_EHIP;
  }
  if (!rc && (hipMemcpy(bp_out, d_bp, n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(obprob_out, d_ob, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(&err, d_err, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess))
    rc = BEEP_EHIP;
  (void)hipFree(d);
  if (!rc && err) rc = BEEP_ENAN;
  return rc;
}

// ---- mixed batches ----

namespace {
// A size class of a mixed batch: cells rounded up to a power of two, the
// compartment structure, the kernel (specialised or generic) and the jobs'
// mode (simulations and MBP proposals run as separate launches, each with the
// kernel compiled for its mode alone, which needs fewer registers).
struct TuneKey {
  int ag2, C, T, spec, mode;
  bool operator<(const TuneKey& o) const {
    return std::tie(ag2, C, T, spec, mode) < std::tie(o.ag2, o.C, o.T, o.spec, o.mode);
  }
};
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tune;
bool g_tune_loaded = false;

const char* tune_path() {
  const char* p = std::getenv("BEEP_TUNE_CACHE");
  return (p && *p) ? p : nullptr;
}

void tune_load() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const char* p = tune_path();
  if (!p) return;
  std::ifstream f(p);
  std::string ver;
  if (!(f >> ver) || ver != "beep-tune-3") return;     // another format: ignored, rewritten on save
  TuneKey k;
  int b;
  while (f >> k.ag2 >> k.C >> k.T >> k.spec >> k.mode >> b)
    if (b >= 64 && b <= (k.spec ? BEEP_SPEC_MAX_BLOCK : BEEP_FUSED_MAX_BLOCK) && b % 64 == 0) g_tune[k] = b;
}

void tune_save() {
  const char* p = tune_path();
  if (!p) return;
  std::ofstream f(p);
  f << "beep-tune-3\n";
  for (auto& kv : g_tune)
    f << kv.first.ag2 << " " << kv.first.C << " " << kv.first.T << " " << kv.first.spec << " " << kv.first.mode << " "
      << kv.second << "\n";
}

int pow2_ceil(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

// The streams a batch launched from c0 spreads its work over (made once).
int ensure_bstreams(beep_ctx* c0) {
  if (c0->bstart) return BEEP_OK;
  HIPCHK(c0, hipEventCreateWithFlags(&c0->bstart, hipEventDisableTiming));
  for (auto& e : c0->bend) HIPCHK(c0, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  for (auto& st : c0->bstream) HIPCHK(c0, hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  return BEEP_OK;
}

// ObservationModel::calculate (src/obsmodel.cc:22-109) of every item's
// instance after a batch: per context one observation-kernel launch over its
// items' instances, the launches spread over the batch streams, one upload of
// the instance lists and one download of the sums.  Each instance's EF is
// kept for the acceptance steps.
// Observation log-likelihoods (ObservationModel::calculate, src/obsmodel.cc:
// 22-35) of the instances inst[k][0..ninst[k]) of contexts ctxs[k], into out
// (the contexts' results one after the other, in that order): each context's
// stored observations in chunks of BEEP_OBS_CHUNK, one workgroup per chunk of
// a group of up to four instances, over the whole device in one launch per
// kernel variant (k_obs_chunks), then each instance's chunk sums added in
// chunk order (k_obs_reduce), which also stores its EF when store_EF.  An
// instance's result depends only on its own trajectory and its context's
// observations.  Staging in c0's scratch, work on c0's stream; with ev0 >= 0
// the kernels are bracketed by c0's timing events ev0 and ev0 + 1.  Returns
// after the results are back.
int obs_loglik_chunked(beep_ctx* c0, const std::vector<beep_ctx*>& ctxs, const std::vector<const int32_t*>& inst,
                       const std::vector<int>& ninst, bool store_EF, double* out, int ev0) {
  hipStream_t st = c0->stream;
  using beepdev::ObsGroup;
  using beepdev::ObsItem;
  constexpr int GN = beepdev::OBS_GROUP;
  size_t N = 0;
  for (int v : ninst) N += (size_t)v;
  if (N == 0) return BEEP_OK;
  // Group and item counts first, to lay out the staging.
  size_t ngroups = 0;
  for (size_t k = 0; k < ctxs.size(); k++) ngroups += ((size_t)ninst[k] + GN - 1) / GN;
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t b_inst = al(N * sizeof(int32_t)), b_groups = al(ngroups * sizeof(ObsGroup)),
               b_items = al(N * sizeof(ObsItem)), b_out = al(N * sizeof(double));
  size_t nparts = 0;
  for (size_t k = 0; k < ctxs.size(); k++) {
    const size_t nchunk = ((size_t)ctxs[k]->obs.nobs + BEEP_OBS_CHUNK - 1) / BEEP_OBS_CHUNK;
    nparts += nchunk * (((size_t)ninst[k] + GN - 1) / GN);
  }
  const size_t b_part = al(nparts * GN * sizeof(double));
  const size_t up = b_inst + b_groups + b_items;
  int rc;
  if ((rc = ensure_scratch(c0, up + b_out + b_part))) return rc;
  char* d = (char*)c0->d_scratch;
  const int32_t* d_inst = (const int32_t*)d;
  ObsGroup* d_groups = (ObsGroup*)(d + b_inst);
  ObsItem* d_items = (ObsItem*)(d + b_inst + b_groups);
  double* d_out = (double*)(d + up);
  double* d_part = (double*)(d + up + b_out);
  std::vector<char> h(up, 0);
  int32_t* h_inst = (int32_t*)h.data();
  ObsItem* h_items = (ObsItem*)(h.data() + b_inst + b_groups);
  // Groups by kernel variant (general / population elements / splines), each
  // variant's groups contiguous and numbered by workgroup from 0.
  std::vector<ObsGroup> byv[8];
  long long wgs[8] = {};
  size_t ii = 0;
  int part = 0;
  for (size_t k = 0; k < ctxs.size(); k++) {
    beep_ctx* c = ctxs[k];
    const int v = (c->obs_general ? 4 : 0) | (c->obs_pop ? 2 : 0) | (c->nospl ? 1 : 0);
    const int nchunk = (int)(((size_t)c->obs.nobs + BEEP_OBS_CHUNK - 1) / BEEP_OBS_CHUNK);
    for (int g0 = 0; g0 < ninst[k]; g0 += GN) {
      ObsGroup G;
      std::memset(&G, 0, sizeof G);
      G.od = c->obs;
      G.tn = c->d_tn;
      G.pt = c->d_pt;
      G.ospl = c->nospl ? c->d_ospl : nullptr;
      G.inst_slot = c->nospl ? c->d_inst_slot : nullptr;
      G.inst = d_inst + ii;
      G.traj_sz = c->traj_sz;
      G.pop_traj_sz = c->pop_traj_sz;
      G.ninst = std::min(GN, ninst[k] - g0);
      G.nchunk = nchunk;
      G.wg0 = (int)wgs[v];
      G.part0 = part;
      for (int q = 0; q < G.ninst; q++) {
        const int32_t in = inst[k][g0 + q];
        h_inst[ii] = in;
        h_items[ii] = ObsItem{part, nchunk, q, store_EF ? c->d_EF + in : nullptr};
        ii++;
      }
      if (nchunk > 0) byv[v].push_back(G);
      wgs[v] += nchunk;
      part += nchunk;
    }
  }
  size_t goff[8], gsum = 0;
  for (int v = 0; v < 8; v++) {
    goff[v] = gsum;
    if (!byv[v].empty())
      std::memcpy(h.data() + b_inst + gsum * sizeof(ObsGroup), byv[v].data(), byv[v].size() * sizeof(ObsGroup));
    gsum += byv[v].size();
    if (wgs[v] > beephost::kGridMaxX) return fail(c0, BEEP_EINVAL, "observation log-likelihood: too many chunks");
  }
  HIPCHK(c0, hipMemcpyAsync(d, h.data(), up, hipMemcpyHostToDevice, st));
  if (ev0 >= 0) record(c0, ev0);
  using KC = decltype(&beepdev::k_obs_chunks<false, false, false>);
  static const KC kc[8] = {beepdev::k_obs_chunks<false, false, false>, beepdev::k_obs_chunks<false, false, true>,
                           beepdev::k_obs_chunks<false, true, false>,  beepdev::k_obs_chunks<false, true, true>,
                           beepdev::k_obs_chunks<true, false, false>,  beepdev::k_obs_chunks<true, false, true>,
                           beepdev::k_obs_chunks<true, true, false>,   beepdev::k_obs_chunks<true, true, true>};
  for (int v = 0; v < 8; v++) {
    if (byv[v].empty()) continue;
    hipLaunchKernelGGL(kc[v], dim3((unsigned)wgs[v]), dim3(256), 0, st, (const ObsGroup*)(d_groups + goff[v]),
                       (int)byv[v].size(), d_part);
    HIPCHK(c0, hipGetLastError());
  }
  hipLaunchKernelGGL(beepdev::k_obs_reduce, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st,
                     (const ObsItem*)d_items, (int)N, (const double*)d_part, d_out);
  HIPCHK(c0, hipGetLastError());
  if (ev0 >= 0) record(c0, ev0 + 1);
  HIPCHK(c0, hipMemcpyAsync(out, d_out, N * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(c0, hipStreamSynchronize(st));
  return BEEP_OK;
}

int batch_loglik(beep_ctx* c0, int n, const beep_work* w, double* L_out) {
  std::vector<beep_ctx*> ctxs;
  std::map<beep_ctx*, std::vector<int>> items;
  for (int i = 0; i < n; i++) {
    auto it = items.find(w[i].ctx);
    if (it == items.end()) {
      ctxs.push_back(w[i].ctx);
      items[w[i].ctx].push_back(i);
    } else {
      it->second.push_back(i);
    }
  }
  std::vector<std::vector<int32_t>> insts(ctxs.size());
  std::vector<const int32_t*> ip(ctxs.size());
  std::vector<int> ni(ctxs.size());
  for (size_t k = 0; k < ctxs.size(); k++) {
    beep_ctx* c = ctxs[k];
    if (!c->d_tn) return fail(c0, BEEP_EINVAL, "BEEP_BATCH_LOGLIK: a context does not keep transnum");
    for (int i : items[c]) insts[k].push_back(w[i].inst);
    ip[k] = insts[k].data();
    ni[k] = (int)insts[k].size();
  }
  int rc;
  for (size_t k = 0; k < ctxs.size(); k++)        // splines' slot lists first (synchronous copies)
    if (ctxs[k]->nospl) {
      for (int i : items[ctxs[k]])
        if (ctxs[k]->inst_slot[w[i].inst] < 0)
          return fail(c0, BEEP_EINVAL, "work %d: no parameter slot for the observation splines", i);
      if ((rc = sync_inst_slot(ctxs[k]))) {
        c0->err = ctxs[k]->err;
        return rc;
      }
    }
  // The kernels are timed on the context stream (timing[1]).
  std::vector<double> h(n);
  if ((rc = obs_loglik_chunked(c0, ctxs, ip, ni, true, h.data(), 2))) return rc;
  c0->timing[1] = elapsed(c0, 2, 3);
  size_t o = 0;
  for (size_t k = 0; k < ctxs.size(); k++)
    for (int i : items[ctxs[k]]) {
      const double L = h[o++];
      if (std::isnan(L)) return fail(c0, BEEP_ENAN, "NaN observation log-likelihood (work %d)", i);
      L_out[i] = L;
    }
  return BEEP_OK;
}
}  // namespace

int beep_run_batch(int n, const beep_work* w, uint64_t seed, uint32_t stream, int flags, double* L_out,
                   double* stats) {
  if (n <= 0 || !w || !w[0].ctx) return BEEP_EINVAL;
  beep_ctx* c0 = w[0].ctx;
  if (stream > 0xFFFF) return fail(c0, BEEP_EINVAL, "stream must be below 65536");
  if ((flags & BEEP_BATCH_LOGLIK) && !L_out) return fail(c0, BEEP_EINVAL, "BEEP_BATCH_LOGLIK needs L_out");
  auto t_start = std::chrono::steady_clock::now();
  // Validate and build jobs.
  struct Item { Job job; beep_ctx* ctx; size_t lds; double cost; int AG; SpecKernels* spec; };
  std::vector<Item> fused;
  double cellsteps = 0;
  // Per context: the instances the batch writes (distinct) and the initial
  // states its MBPs read (none of them written by the batch).
  {
    std::map<beep_ctx*, std::pair<std::vector<int32_t>, std::vector<int32_t>>> wr;
    for (int i = 0; i < n; i++) {
      beep_ctx* c = w[i].ctx;
      if (!c) return fail(c0, BEEP_EINVAL, "work %d has no context", i);
      if (c->device != c0->device) return fail(c0, BEEP_EINVAL, "batch contexts must share one device");
      // Step-wise contexts (dense mixing, or an instance beyond LDS) are not
      // batched: their work goes through beep_simulate / beep_mbp.
      if (c->stepwise)
        return fail(c0, BEEP_EINVAL, "work %d: a step-wise context (dense mixing or an instance beyond LDS) cannot be "
                                     "batched; call beep_simulate / beep_mbp on it", i);
      auto& e = wr[c];
      e.first.push_back(w[i].inst);
      if (w[i].kind == BEEP_WORK_MBP) e.second.push_back(w[i].init);
    }
    for (auto& kv : wr) {
      beep_ctx* c = kv.first;
      const auto& written = kv.second.first;
      const auto& read = kv.second.second;
      int rc0 = check_instances(c, (int)written.size(), written.data(), nullptr, "beep_run_batch");
      if (!rc0) {
        std::vector<char> isw(c->capacity, 0);
        for (int32_t x : written) isw[x] = 1;
        for (int32_t x : read)
          if (!valid_inst(c, x) || isw[x]) {
            rc0 = fail(c, BEEP_EINVAL, "beep_run_batch: MBP initial instance %d out of range or written by the batch", x);
            break;
          }
      }
      if (rc0) {
        if (c != c0) c0->err = c->err;
        return rc0;
      }
    }
  }
  for (int i = 0; i < n; i++) {
    beep_ctx* c = w[i].ctx;
    const auto& m = c->m;
    if (!valid_slot(c, w[i].slot) || !c->slot_set[w[i].slot])
      return fail(c0, BEEP_EINVAL, "work %d: bad slot", i);
    Job j;
    if (w[i].kind == BEEP_WORK_SIMULATE) {
      if (w[i].t0 < 0 || w[i].t1 > m.S || w[i].t0 > w[i].t1) return fail(c0, BEEP_EINVAL, "work %d: bad steps", i);
      j = make_job(c, w[i].inst, w[i].slot, w[i].key, w[i].t0, w[i].t1);
    } else if (w[i].kind == BEEP_WORK_MBP) {
      if (!c->d_tn || !c->d_tm || !valid_inst(c, w[i].init) || w[i].init == w[i].inst)
        return fail(c0, BEEP_EINVAL, "work %d: bad MBP pair", i);
      j = make_job(c, w[i].inst, w[i].slot, w[i].key, 0, m.S);
      j.mode = beepdev::MODE_MBP;
      j.init_tn = c->d_tn + (size_t)w[i].init * c->traj_sz;
      j.init_tm = c->d_tm + (size_t)w[i].init * c->traj_sz;
    } else {
      return fail(c0, BEEP_EINVAL, "work %d: unknown kind", i);
    }
    cellsteps += (double)m.AG * (j.t1 - j.t0);
    j.err = c0->d_err;                       // one error word for the whole batch
    if (j.t0 == j.t1) continue;
    SpecKernels* sk = c->spec;
    size_t lds = sk ? c->spec_lds : beepdev::fused_lds_bytes(m.A, m.G, m.C, m.T, c->h_topo.nnz, c->h_topo.stage_mix);
    fused.push_back({j, c, lds, (double)m.AG * m.T * (j.t1 - j.t0), m.AG, sk});
  }
  int rc;
  int nlaunch = 0;
  const int nstep = 0;
  if (!fused.empty()) {
    // Size classes (and kernel: one specialised kernel per compartment
    // structure; and mode), each sorted longest first.
    std::map<std::pair<TuneKey, SpecKernels*>, std::vector<Item>> groups;
    for (auto& it : fused)
      groups[{{pow2_ceil(it.AG), it.ctx->m.C, it.ctx->m.T, it.spec ? 1 : 0, it.job.mode}, it.spec}].push_back(it);
    std::vector<Job> all;
    // key, offset, count, lds, maxAG, kernel, every job starts from the initial state, the longest job's cost
    std::vector<std::tuple<TuneKey, size_t, size_t, size_t, int, SpecKernels*, bool, double>> launches;
    for (auto& kv : groups) {
      auto& v = kv.second;
      std::stable_sort(v.begin(), v.end(), [](const Item& a, const Item& b) { return a.cost > b.cost; });
      size_t lds = 0;
      int maxAG = 0;
      bool rerunnable = true;
      for (auto& it : v) {
        lds = std::max(lds, it.lds);
        maxAG = std::max(maxAG, it.AG);
        rerunnable = rerunnable && it.job.from_init;
      }
      // A coupled proposal costs about 1.7 times a simulation per cell-step.
      const double lead = v.front().cost * (kv.first.first.mode == beepdev::MODE_MBP ? 1.7 : 1.0);
      launches.emplace_back(kv.first.first, all.size(), v.size(), lds, maxAG, kv.first.second, rerunnable, lead);
      for (auto& it : v) all.push_back(it.job);
    }
    // Launches in order of their longest job, longest first, dealt round-robin
    // over the batch streams: the classes whose workgroups run longest start
    // first, and the short ones fill the device as those drain.
    std::stable_sort(launches.begin(), launches.end(),
                     [](const auto& a, const auto& b) { return std::get<7>(a) > std::get<7>(b); });
    if (all.size() > c0->bjobs_cap) {
      if (c0->d_bjobs) {
        HIPCHK(c0, hipStreamSynchronize(c0->stream));
        HIPCHK(c0, hipFree(c0->d_bjobs));
        c0->allocs.erase(std::find(c0->allocs.begin(), c0->allocs.end(), (void*)c0->d_bjobs));
        c0->allocated -= c0->bjobs_cap * sizeof(Job);
        c0->d_bjobs = nullptr;
        c0->bjobs_cap = 0;
      }
      size_t cap = std::max<size_t>(all.size(), 1024);
      if ((rc = dalloc(c0, (void**)&c0->d_bjobs, cap * sizeof(Job)))) return rc;
      c0->bjobs_cap = cap;
    }
    HIPCHK(c0, hipMemcpyAsync(c0->d_bjobs, all.data(), all.size() * sizeof(Job), hipMemcpyHostToDevice, c0->stream));
    // The size classes run on a few streams, so that the last, longest
    // instances of one launch overlap the next launch instead of leaving the
    // device idle; each stream starts after the job upload.  (Autotuning
    // launches run alone on the context stream to be timed.)
    const int nbs = (int)(sizeof c0->bstream / sizeof c0->bstream[0]);
    if ((rc = ensure_bstreams(c0))) return rc;
    // The fused phase is timed on the context stream from before the batch
    // streams start to after they have all ended (timing[0]).
    record(c0, 0);
    HIPCHK(c0, hipEventRecord(c0->bstart, c0->stream));
    for (auto& st : c0->bstream) HIPCHK(c0, hipStreamWaitEvent(st, c0->bstart, 0));
    int next_stream = 0;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load();
    bool tuned_any = false;
    for (auto& L : launches) {
      const TuneKey key = std::get<0>(L);
      size_t off = std::get<1>(L), cnt = std::get<2>(L), lds = std::get<3>(L);
      int maxAG = std::get<4>(L);
      SpecKernels* sk = std::get<5>(L);
      const bool rerunnable = std::get<6>(L);
      const bool sim = key.mode == beepdev::MODE_SIM;
      auto launch_any = [&](int blk, hipStream_t st) {
        const Job* dj = c0->d_bjobs + off;
        if (sk) return launch_spec(c0, sk, key.mode, dj, (int)cnt, lds, blk, seed, stream, st);
        return sim ? launch_fused_t<beepdev::MODE_SIM>(c0, dj, (int)cnt, lds, blk, seed, stream, st)
                   : launch_fused_t<beepdev::MODE_MBP>(c0, dj, (int)cnt, lds, blk, seed, stream, st);
      };
      int block = sk ? spec_block(maxAG) : fused_block(maxAG);
      auto f = g_tune.find(key);
      if (f != g_tune.end()) block = f->second;
      else if ((flags & BEEP_BATCH_AUTOTUNE) && rerunnable) {
        // Same keys give the same trajectories whatever the workgroup size,
        // and every job of the class starts from the initial state (t0 = 0:
        // simulations from the start, MBPs), so each candidate run recomputes
        // the same results from the same inputs; the fastest is kept.  A class
        // with a job that continues from its current state (t0 > 0) would
        // advance that state once per candidate: it is not tuned here and
        // runs once with the default size.
        float best = 1e30f;
        const int maxb = sk ? BEEP_SPEC_MAX_BLOCK : BEEP_FUSED_MAX_BLOCK;
        for (int cand = 64; cand <= maxb; cand *= 2) {
          if (cand > pow2_ceil((maxAG + 63) / 64 * 64)) break;
          record(c0, 0);
          if ((rc = launch_any(cand, nullptr))) return rc;
          record(c0, 1);
          HIPCHK(c0, hipStreamSynchronize(c0->stream));
          float ms = (float)elapsed(c0, 0, 1);
          nlaunch++;
          if (ms < best) { best = ms; block = cand; }
        }
        int fb = fused_block(maxAG);
        if (!sk && fb > block && (fb & (fb - 1))) {   // the one-pass size when it is not a power of two
          record(c0, 0);
          if ((rc = launch_any(fb, nullptr))) return rc;
          record(c0, 1);
          HIPCHK(c0, hipStreamSynchronize(c0->stream));
          float ms = (float)elapsed(c0, 0, 1);
          nlaunch++;
          if (ms < best) { best = ms; block = fb; }
        }
        g_tune[key] = block;
        tuned_any = true;
        continue;
      }
      // An autotuning call keeps every launch on the context stream, so that
      // the timed candidates run alone.
      hipStream_t st = (flags & BEEP_BATCH_AUTOTUNE) ? nullptr : c0->bstream[next_stream++ % nbs];
      if ((rc = launch_any(block, st))) return rc;
      nlaunch++;
    }
    if (tuned_any) tune_save();
    // Join: the context stream waits for every batch stream, then records the
    // end of the fused phase.
    for (int k = 0; k < nbs; k++) {
      HIPCHK(c0, hipEventRecord(c0->bend[k], c0->bstream[k]));
      HIPCHK(c0, hipStreamWaitEvent(c0->stream, c0->bend[k], 0));
    }
    record(c0, 1);
    HIPCHK(c0, hipStreamSynchronize(c0->stream));   // `all` is the source of an async copy
    c0->timing[0] = (flags & BEEP_BATCH_AUTOTUNE) ? -1 : elapsed(c0, 0, 1);
    c0->timing[3] = nlaunch;
  }
  HIPCHK(c0, hipStreamSynchronize(c0->stream));   // also keeps `all` alive until its copy is done
  // Errors raised by the batch's kernels: every job reported into the first
  // context's error word.
  if ((rc = check_device_error(c0))) return rc;
  if (flags & BEEP_BATCH_LOGLIK) {
    if ((rc = batch_loglik(c0, n, w, L_out))) return rc;
  }
  if (stats) {
    stats[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    stats[1] = nlaunch;
    stats[2] = nstep;
    stats[3] = cellsteps;
  }
  return BEEP_OK;
}

int beep_batch_accept(int n, const beep_work* w, int mode, const double* invT, double EFcut, const double* dlogq,
                      uint64_t seed, uint32_t stream, double* al_out, uint8_t* accepted) {
  if (n <= 0 || !w || !w[0].ctx) return BEEP_EINVAL;
  beep_ctx* c0 = w[0].ctx;
  if (!invT || !dlogq || !al_out || !accepted || (mode != 0 && mode != 1) || stream > 0xFFFF)
    return fail(c0, BEEP_EINVAL, "beep_batch_accept: bad arguments");
  // Packed upload: EF pointers of the proposed and initial instances, the
  // temperatures, the log proposal ratios and the keys.
  std::vector<char> h((size_t)n * (8 + 8 + 8 + 8 + 4));
  const double** efp = (const double**)h.data();
  const double** efi = efp + n;
  double* hT = (double*)(efi + n);
  double* hq = hT + n;
  uint32_t* hk = (uint32_t*)(hq + n);
  for (int i = 0; i < n; i++) {
    beep_ctx* c = w[i].ctx;
    if (!c || c->device != c0->device) return fail(c0, BEEP_EINVAL, "work %d: no context or another device", i);
    if (w[i].kind != BEEP_WORK_MBP || !valid_inst(c, w[i].inst) || !valid_inst(c, w[i].init))
      return fail(c0, BEEP_EINVAL, "work %d: not an MBP proposal with valid instances", i);
    efp[i] = c->d_EF + w[i].inst;
    efi[i] = c->d_EF + w[i].init;
    hT[i] = invT[i];
    hq[i] = dlogq[i];
    hk[i] = w[i].key;
  }
  int rc;
  const size_t inb = (h.size() + 255) / 256 * 256;
  if ((rc = ensure_scratch(c0, inb + (size_t)n * 9))) return rc;
  char* d = (char*)c0->d_scratch;
  double* d_al = (double*)(d + inb);
  uint8_t* d_acc = (uint8_t*)(d_al + n);
  HIPCHK(c0, hipMemcpyAsync(d, h.data(), h.size(), hipMemcpyHostToDevice, c0->stream));
  hipLaunchKernelGGL(beepdev::k_accept_ptrs, dim3((n + 255) / 256), dim3(256), 0, c0->stream, n,
                     (const double* const*)d, (const double* const*)(d + (size_t)8 * n), (const double*)(d + (size_t)16 * n),
                     (const double*)(d + (size_t)24 * n), (const uint32_t*)(d + (size_t)32 * n), mode, EFcut, seed, stream,
                     d_al, d_acc);
  HIPCHK(c0, hipGetLastError());
  HIPCHK(c0, hipMemcpyAsync(al_out, d_al, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c0->stream));
  HIPCHK(c0, hipMemcpyAsync(accepted, d_acc, (size_t)n, hipMemcpyDeviceToHost, c0->stream));
  HIPCHK(c0, hipStreamSynchronize(c0->stream));
  for (int i = 0; i < n; i++)
    if (std::isnan(al_out[i])) return fail(c0, BEEP_ENAN, "NaN acceptance probability (work %d)", i);   // src/mbp.cc:353
  return BEEP_OK;
}

int beep_batch_mc3_swap(int nladder, const int32_t* ladder_ptr, beep_ctx* const* ctx, const int32_t* inst,
                        const int32_t* run, const double* invT, uint64_t seed, uint32_t stream, int32_t* order_out,
                        int32_t* ntr_out, int32_t* nac_out) {
  if (nladder <= 0 || !ladder_ptr || !ctx || !ctx[0] || !inst || !run || !invT || !order_out || !ntr_out || !nac_out)
    return BEEP_EINVAL;
  beep_ctx* c0 = ctx[0];
  if (stream > 0xFFFF || ladder_ptr[0] != 0) return fail(c0, BEEP_EINVAL, "beep_batch_mc3_swap: bad arguments");
  for (int l = 0; l < nladder; l++)
    if (ladder_ptr[l + 1] <= ladder_ptr[l]) return fail(c0, BEEP_EINVAL, "ladder %d is empty or out of order", l);
  const int n = ladder_ptr[nladder];
  // Packed upload: ladder offsets, EF pointers, temperatures, run labels.
  std::vector<char> h((size_t)(nladder + 1) * 4 + (size_t)n * (8 + 8 + 4) + 16);
  int* hp = (int*)h.data();
  const size_t off_ef = ((size_t)(nladder + 1) * 4 + 7) / 8 * 8;
  const double** hef = (const double**)(h.data() + off_ef);
  double* hT = (double*)(hef + n);
  int* hr = (int*)(hT + n);
  for (int l = 0; l <= nladder; l++) hp[l] = ladder_ptr[l];
  for (int i = 0; i < n; i++) {
    beep_ctx* c = ctx[i];
    if (!c || c->device != c0->device || !valid_inst(c, inst[i]))
      return fail(c0, BEEP_EINVAL, "chain %d: no context, another device or instance out of range", i);
    hef[i] = c->d_EF + inst[i];
    hT[i] = invT[i];
    hr[i] = run[i];
  }
  int rc;
  const size_t inb = (h.size() + 255) / 256 * 256;
  if ((rc = ensure_scratch(c0, inb + (size_t)n * (8 + 3 * 4)))) return rc;
  char* d = (char*)c0->d_scratch;
  double* d_EF = (double*)(d + inb);
  int* d_order = (int*)(d_EF + n);
  int* d_ntr = d_order + n;
  int* d_nac = d_ntr + n;
  HIPCHK(c0, hipMemcpyAsync(d, h.data(), h.size(), hipMemcpyHostToDevice, c0->stream));
  hipLaunchKernelGGL(beepdev::k_mc3_swap_ladders, dim3((nladder + 63) / 64), dim3(64), 0, c0->stream, nladder,
                     (const int*)d, (const int*)(d + off_ef + (size_t)16 * n), (const double*)(d + off_ef + (size_t)8 * n),
                     (const double* const*)(d + off_ef), d_EF, d_order, d_ntr, d_nac, seed, stream);
  HIPCHK(c0, hipGetLastError());
  HIPCHK(c0, hipMemcpyAsync(order_out, d_order, (size_t)n * 4, hipMemcpyDeviceToHost, c0->stream));
  HIPCHK(c0, hipMemcpyAsync(ntr_out, d_ntr, (size_t)n * 4, hipMemcpyDeviceToHost, c0->stream));
  HIPCHK(c0, hipMemcpyAsync(nac_out, d_nac, (size_t)n * 4, hipMemcpyDeviceToHost, c0->stream));
  HIPCHK(c0, hipStreamSynchronize(c0->stream));
  return BEEP_OK;
}

int beep_synchronize(beep_ctx* c) {
  if (!c) return BEEP_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BEEP_OK;
}

// ---- diagnostics ----


int beep_debug_philox(int device, int n, const uint32_t* ctr, uint32_t k0, uint32_t k1, uint32_t* out) {
  if (n <= 0 || !ctr || !out) return BEEP_EINVAL;
  return debug_run(device, (size_t)n * 16, ctr, 0, nullptr, (size_t)n * 16, out, [&](void* a, void*, void* o) {
    hipLaunchKernelGGL(beepdev::k_dbg_philox, dim3((n + 255) / 256), dim3(256), 0, 0, n, (const beepdev::u4*)a, k0, k1,
                       (beepdev::u4*)o);
  });
}

int beep_debug_poisson(int device, int n, const float* lam, uint64_t seed, uint32_t key, int32_t* out) {
  if (n <= 0 || !lam || !out) return BEEP_EINVAL;
  for (int i = 0; i < n; i++)
    if (!(lam[i] >= 0.f) || lam[i] > 1e9f) return BEEP_EINVAL;
  return debug_run(device, (size_t)n * 4, lam, 0, nullptr, (size_t)n * 4, out, [&](void* a, void*, void* o) {
    hipLaunchKernelGGL(beepdev::k_dbg_poisson, dim3((n + 255) / 256), dim3(256), 0, 0, n, (const float*)a, seed, key,
                       (int*)o);
  });
}

int beep_debug_gemm(int device, int rows, int K, const float* A, const float* B, float* C) {
  if (rows <= 0 || K <= 0 || K % 64 || !A || !B || !C) return BEEP_EINVAL;
  const long long tiles = (long long)((rows + 63) / 64) * (K / 64);
  if (tiles > beephost::kGridMaxX / 2) return BEEP_EINVAL;
  // The operands as the dense contraction holds them: A over rows padded to
  // 64, B transposed, both in fragment order.
  const size_t rows_pad = ((size_t)rows + 63) / 64 * 64;
  std::vector<float> Af(rows_pad * K, 0.f), Btf((size_t)K * K);
  for (int r = 0; r < rows; r++)
    for (int k = 0; k < K; k++) Af[beepdev::frag_at(r, k, K)] = A[(size_t)r * K + k];
  for (int k = 0; k < K; k++)
    for (int j = 0; j < K; j++) Btf[beepdev::frag_at(j, k, K)] = B[(size_t)k * K + j];
  return debug_run(device, Af.size() * 4, Af.data(), Btf.size() * 4, Btf.data(), (size_t)rows * K * 4, C,
                   [&](void* a, void* b, void* o) {
                     hipLaunchKernelGGL(beepdev::k_dbg_gemm, dim3((unsigned)((tiles + 7) / 8 * 8)), dim3(256), 0, 0,
                                        (const float*)a, (const float*)b, (float*)o, rows, K);
                   });
}

int beep_debug_check_launch(int kind, const int64_t* args, int nargs, int64_t* grid_out, char* err, int errlen) {
  std::string e;
  if (!args) {
    e = "no arguments";
  } else if (kind == BEEP_CHECK_FUSED && nargs == 4) {
    beephost::Dim g;
    e = beephost::fused_geometry(args[0], (int)args[1], (size_t)args[2], (size_t)args[3], BEEP_FUSED_MAX_BLOCK, &g);
    if (e.empty() && grid_out) grid_out[0] = g.x;
  } else if (kind == BEEP_CHECK_STEPWISE && nargs == 5) {
    beephost::StepGeometry g;
    e = beephost::stepwise_geometry((int)args[0], (int)args[1], (int)args[2], args[3], args[4], BEEP_STEP_BLOCK, &g);
    if (e.empty() && grid_out) {
      const long long v[6] = {g.cell.x, g.cell.y, g.dense.x, g.dense.y, g.sparse.x, g.sparse.y};
      for (int k = 0; k < 6; k++) grid_out[k] = v[k];
    }
  } else {
    e = "unknown check or wrong argument count";
  }
  if (e.empty()) return BEEP_OK;
  if (err && errlen > 0) {
    std::strncpy(err, e.c_str(), errlen - 1);
    err[errlen - 1] = 0;
  }
  return BEEP_EINVAL;
}

int beep_debug_binomial(int device, int n, const int32_t* nn, const double* p, uint64_t seed, uint32_t key,
                        int32_t* out) {
  if (n <= 0 || !nn || !p || !out) return BEEP_EINVAL;
  for (int i = 0; i < n; i++)
    if (nn[i] < 0 || !(p[i] > 0.0) || !(p[i] < 1.0)) return BEEP_EINVAL;
  return debug_run(device, (size_t)n * 4, nn, (size_t)n * 8, p, (size_t)n * 4, out, [&](void* a, void* b, void* o) {
    hipLaunchKernelGGL(beepdev::k_dbg_binomial, dim3((n + 255) / 256), dim3(256), 0, 0, n, (const int*)a,
                       (const double*)b, seed, key, (int*)o);
  });
}

}  // extern "C"
