// Host-visible launch descriptors of the force-field kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mm {

constexpr int kCellBlock = 256;
// Cell passes of more blocks than this run their two instances concurrently; smaller ones run
// both kinds of block in one launch (cell_kernel_both).
constexpr int kForkMinBlocks = 512;
// Node gathers over at most this many 256-node blocks (at most four waves a SIMD) load all of
// a node's records at once (sum_node_records<true>): latency bounds such a pass. NVT-NHC / NPT
// step, records one at a time -> all at once: 32^3 46.3 / 105.9 -> 43.9 / 103.4 us, 48^3
// 57.5 / 126.5 -> 54.1 / 121.6 us, 64^3 74.4 / 170.4 -> 73.3 / 170.9 us; with more blocks the
// extra registers cost more than the loads in flight gain (80^3 NPT 223 -> 234 us, 128^3
// 650 -> 666 us). The same bits either way.
constexpr int kGatherAllLoadsMaxBlocks = 1024;
constexpr int kPartialStride = 16;
constexpr int kMaxStates = 8;  // metastable states per cell type

struct ForceArgs {
  const double* pos;       // [N][3]
  const double* rvecs;     // [3][3] device (of system 0; system s at rvecs + s * rvecs_stride)
  int64_t rvecs_stride = 0;
  const int32_t* blk_sys = nullptr;  // [W / kCellBlock] system of each block (null: all 0)
  const int32_t* cnode;    // [W][8] corner nodes in work order (-1 = padding slot)
  const int32_t* cwrap;    // [W] wrap bits of each work slot
  const int32_t* blk_row;  // [W / kCellBlock] table row of every block
  const double* tab;       // [rows][max_states][kTabStride]
  const double* kt;        // [rows] boltzmann * effective temperature
  const int32_t* nstates;  // [rows]
  int32_t max_states;
  int64_t W;               // work slots (multiple of kCellBlock)
  int64_t N;
  const int32_t* ncell;    // [8][N] node -> record e (-1 = none): slot e >> rshift, class e & mask
  double* gcell;           // component j of class p of slot s at (3 p + j) * rstride + s
  int64_t rstride = 0;     // extent of a gradient plane (slots plus compact upper faces)
  int rshift = 2;          // 2: four face classes (the cell pass); 3: eight corners (the adjoint)
  const int32_t* cbase = nullptr;  // [W / 64] first compact upper-face slot of every wave
  double* partials;        // [kPartialStride][W / kCellBlock] (SoA)
  double* energy;          // [1] or null
  double* gpos;            // [N][3] or null
  double* vtens;           // [9] or null
  int accumulate;
  int model;               // 0 corner8, 1 original
  int* nanflag;
  int force_grad = 0;       // compute gradients even when gpos and vtens are null
  int corner_virial = 0;    // 1: the virial as the corners' sum of B^T S B (no vertex gradient
                            //    needed for it)
  int store_records = 1;    // with corner_virial: 0 forms no gradient and stores no records
  // pending position transforms [3][3] (row vectors, system s at xmat + s * rvecs_stride) that
  // the pass applies to every node position it reads, xmat2 first when the flag at
  // xmat2_on + s * rvecs_stride is set (both null: none): the position scaling of a barostat
  // call that its own pass does not store. Only for the passes with gradients and the vertex
  // virial, or with the corner virial and no gradients.
  const double* xmat = nullptr;
  const double* xmat2 = nullptr;
  const double* xmat2_on = nullptr;
  int split = -1;           // multistate blocks are [0, split), single-state [split, nb); -1: mixed
  hipStream_t side = nullptr;   // stream of the multistate instance (null: the caller's stream)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

// Parameter table [nrows][max_states][kTabStride] and kt [nrows] from device arrays (h0 [.][9],
// Voigt cv [.][36], free_energy, temp_eff; nstates device), as the host builder computes them.
hipError_t launch_build_table(int corner8, int nrows, int max_states, const int32_t* nstates,
                              const double* h0, const double* cv, const double* free_energy,
                              const double* temp_eff, double boltzmann, double* tab, double* kt,
                              hipStream_t stream);

// Reductions of device arrays into single values, one workgroup per item (reduce_items_kernel).
constexpr int kReduceSum = 0;
constexpr int kReduceMax = 1;
constexpr int kReduceItems = 16;
struct ReduceItem {
  const double* src;
  int64_t n;
  double* dst;
  int op;      // kReduceSum (fixed order) or kReduceMax (NaN-propagating)
  int acc;     // add the total to dst[0] instead of overwriting it
  int nanbit;  // bit set in the NaN flag when the total is NaN (0: none)
};
struct ReduceList {
  ReduceItem item[kReduceItems];
  int count = 0;
  void add(const double* src, int64_t n, int op, double* dst, int acc, int nanbit) {
    item[count++] = ReduceItem{src, n, dst, op, acc, nanbit};
  }
};
hipError_t launch_reduce(const ReduceList& list, int* nanflag, hipStream_t stream);

hipError_t launch_force(const ForceArgs& a, hipStream_t stream);
// Energy (partial row 0) and virial (rows 1-9) items of a single-system cell pass with nb blocks.
void add_force_totals(ReduceList& list, const ForceArgs& a, int nb);
// Node gather of a line-search trial: also the per-workgroup partials of sum_n gpos_n . dir_n,
// sum_n |gpos_n|^2, sum_n gpos_n . gold_n and max_n |gpos_n|^2, rows 0-3 of dpart (row stride
// (N + 255) / 256).
hipError_t launch_gather_dot(const ForceArgs& a, const double* dir, const double* gold,
                             double* dpart, hipStream_t stream);
// Line-search trial point over N nodes: d <- beta d - g when set_dir, then xa <- x + alpha d;
// per-workgroup partials of sum_n |xa_n - x_n|^2 and max_n |xa_n - x_n|^2, rows 0-1 of part
// (row stride (N + 255) / 256).
hipError_t launch_cg_trial(const double* x, const double* g, double* d, double beta, int set_dir,
                           double alpha, double* xa, int64_t N, double* part,
                           hipStream_t stream);
// out[0..5) = sum_n |g_n|^2, sum_n g_n . go_n, sum_n |x_n - xo_n|^2, max_n |g_n|^2,
// max_n |x_n - xo_n|^2 over N nodes of 3 values; part holds 5 (N + 255) / 256 doubles.
hipError_t launch_cg_stats(const double* x, const double* xo, const double* g, const double* go,
                           int64_t N, double* part, double* out, int* nanflag,
                           hipStream_t stream);
// Per-system energy and virial of a batch: energies[s], vtens[9 s] (either may be null).
hipError_t launch_batch_finalize(const double* partials, int nb, const int32_t* sys_ptr,
                                 const int32_t* sys_idx, int nsys, double* energies,
                                 double* vtens, int* nanflag, hipStream_t stream);
// Node gather of gcell into gpos (overwrite or accumulate, as a.accumulate).
hipError_t launch_gather(const ForceArgs& a, hipStream_t stream);
// Cell kernel only: per-cell gradients into gcell, energy/virial block partials.
hipError_t launch_cells(const ForceArgs& a, hipStream_t stream);

}  // namespace mm
