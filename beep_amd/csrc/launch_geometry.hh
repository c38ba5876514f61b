// Launch geometry of the simulation kernels, computed and checked on the host
// before anything is enqueued (pure functions: no device calls, so the checks
// run and are tested without a GPU).
//
// Every launch of the fused and step-wise paths takes its grid and workgroup
// from here; a geometry that a kernel's indexing or its buffers do not allow
// is refused with a message instead of being handed to the device.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>

namespace beephost {

// Grid limits of a dispatch on gfx950: x up to 2^31 - 1 workgroups, y and z
// up to 65535, and the work-item count of each dimension inside 32 bits.
constexpr long long kGridMaxX = 2147483647LL;
constexpr long long kGridMaxYZ = 65535LL;

struct Dim { long long x = 1, y = 1, z = 1; };

inline std::string check_grid(const char* what, Dim grid, Dim block) {
  if (grid.x < 1 || grid.y < 1 || grid.z < 1) return std::string(what) + ": empty grid";
  if (grid.x > kGridMaxX || grid.y > kGridMaxYZ || grid.z > kGridMaxYZ)
    return std::string(what) + ": grid beyond the dispatch limits (x <= 2^31-1, y, z <= 65535)";
  if (grid.x * block.x > 0xFFFFFFFFLL || grid.y * block.y > 0xFFFFFFFFLL || grid.z * block.z > 0xFFFFFFFFLL)
    return std::string(what) + ": work-item count of a dimension beyond 32 bits";
  if (block.x < 1 || block.y < 1 || block.z < 1 || block.x * block.y * block.z > 1024)
    return std::string(what) + ": workgroup size outside 1..1024";
  return "";
}

// One launch of a fused kernel: one workgroup per job, `block` threads (whole
// wavefronts, at most max_block, the kernel's launch bound), `lds` bytes of
// dynamic LDS against the kernel's limit.
inline std::string fused_geometry(long long njobs, int block, size_t lds, size_t lds_limit, int max_block,
                                  Dim* grid_out = nullptr, Dim* block_out = nullptr) {
  if (njobs < 1) return "fused launch: no jobs";
  if (block < 64 || block > max_block || block % 64) return "fused launch: workgroup size not whole wavefronts within the kernel's bound";
  if (lds > lds_limit) return "fused launch: dynamic LDS above the kernel's limit";
  Dim g, b;
  g.x = njobs;
  b.x = block;
  std::string e = check_grid("fused launch", g, b);
  if (!e.empty()) return e;
  if (grid_out) *grid_out = g;
  if (block_out) *block_out = b;
  return "";
}

// The step-wise path's launches for n jobs of one context (A areas, G age
// groups, Apad the padded area count of its D / I buffers, which hold
// capacity * G rows): the initial infectivity and transition kernels (grid
// n x ceil(AG / step_block): job on x, block of cells on y), the dense MFMA
// contraction (its ceil(rows/64) x Apad/64 tiles of 64x64 on a one-dimensional
// grid rounded up to a multiple of the eight XCDs, tiles dealt to XCDs in
// contiguous runs) or the sparse one (rows x ceil(A/256)).
struct StepGeometry {
  Dim cell;       // k_step_init and the transition kernels
  Dim dense;      // k_dense_contract
  Dim sparse;     // k_sparse_contract
  long long rows = 0;
};

inline std::string stepwise_geometry(int A, int G, int Apad, long long n, long long capacity, int step_block,
                                     StepGeometry* out) {
  if (A < 1 || G < 1) return "step-wise launch: empty model";
  if (n < 1) return "step-wise launch: no jobs";
  if (n > capacity) return "step-wise launch: more jobs than the context's instances (the infectivity buffers hold capacity rows)";
  if (Apad < A || Apad % 64) return "step-wise launch: padded area count not a multiple of 64 covering the areas";
  if (step_block < 64 || step_block % 64) return "step-wise launch: bad cells per workgroup";
  StepGeometry g;
  g.rows = n * G;
  g.cell.x = n;
  g.cell.y = ((long long)A * G + step_block - 1) / step_block;
  const long long tiles = ((g.rows + 63) / 64) * (Apad / 64);
  g.dense.x = (tiles + 7) / 8 * 8;
  g.dense.y = 1;
  g.sparse.x = g.rows;
  g.sparse.y = (A + 255) / 256;
  Dim b256;
  b256.x = 256;
  Dim bstep;
  bstep.x = step_block;
  std::string e = check_grid("step-wise cell kernels", g.cell, bstep);
  if (e.empty()) e = check_grid("dense contraction", g.dense, b256);
  if (e.empty()) e = check_grid("sparse contraction", g.sparse, b256);
  if (!e.empty()) return e;
  if (out) *out = g;
  return "";
}

}  // namespace beephost
