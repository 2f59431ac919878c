// Host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace mm {

int set_error(int code, const char* fmt, ...);
int hip_error(hipError_t e, const char* what);
double inverse3(const double* h, double* inv);
void voigt_average(const double* C, double* Cv);
int build_param_table(int model, int32_t nrows, int32_t max_states, const int32_t* nstates,
                      const double* h0, const double* c0, const double* free_energy,
                      const double* temp_eff, double boltzmann, std::vector<double>& tab,
                      std::vector<double>& kt, bool c_is_voigt);

}  // namespace mm

// Device-resident force field (include/micmec_amd.h).
struct mm_field {
  int device = 0;
  int model = 0;
  int64_t C = 0, N = 0, W = 0;
  int32_t nrows = 0, max_states = 0;
  int32_t* cnode = nullptr;
  int32_t* cwrap = nullptr;
  int32_t* blk_row = nullptr;
  int32_t* ncell = nullptr;           // [8][N] node -> 8 w + k of its corners (adjoint layout)
  int32_t* nrec = nullptr;            // [8][N] node -> face-class records of the cell pass
  int32_t* cbase = nullptr;           // [W/64] see ForceArgs::cbase
  int64_t S = 0;                      // extent of a gradient plane of the cell pass
  int32_t* nstates = nullptr;
  double* tab = nullptr;
  double* kt = nullptr;
  double* gcell = nullptr;
  double* partials = nullptr;
  double* cgpart = nullptr;           // [8][(N + 255) / 256] node-block partials (mm_cg_*)
  double* cghost = nullptr;           // [8] pinned host staging of mm_cg_* results
  int* nanflag = nullptr;
  int32_t nsys = 1;                   // independent systems sharing the object
  int32_t* blk_sys = nullptr;         // [W/256] system of every cell block
  int32_t* sys_cblk_ptr = nullptr;    // [nsys + 1] CSR of the cell blocks of every system
  int32_t* sys_cblk_idx = nullptr;    // [W/256]
  std::vector<int64_t> slot_of_cell;  // reference cell index -> work slot
  std::vector<int32_t> blk_row_host;  // [W/256] copy of blk_row
  std::vector<int32_t> nstates_host;  // [nrows] copy of nstates
  int split = -1;                     // see ForceArgs::split
  hipStream_t side = nullptr;         // multistate instance of the cell pass runs here
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

namespace mm {
// Recompute mm_field::split from the state counts of the rows (host array).
void update_split(mm_field* f, const int32_t* nstates);
// Fill the cell-pass scheduling fields of a ForceArgs from the field.
struct ForceArgs;
void schedule_args(const mm_field* f, ForceArgs& a);
}  // namespace mm

