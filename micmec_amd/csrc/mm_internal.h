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

// Spatial tiles of the cell pass (mm_tile.hip): every workgroup owns a box of cells, stages
// the positions of the box's node slots in LDS and sums the vertex gradients of its cells per
// slot there; a node then sums its slots over the tiles that hold it (CSR hptr/hent).
struct mm_tiles {
  int enabled = 0;
  int ntiles = 0;
  int tx = 0, ty = 0, tz = 0;  // cells per tile along x, y, z
  int S = 0;                   // node slots per tile, (tx + 1)(ty + 1)(tz + 1)
  int maxw = 0;                // waves per tile (lanes sorted by parameter row, padded to 64)
  int64_t E = 0;               // CSR entries
  int4* meta = nullptr;        // [ntiles] origin x, y, z, system
  int4* geom = nullptr;        // [nsys] cells along x, y, z, periodic (0/1)
  int32_t* node = nullptr;     // [ntiles][S] node of every slot (-1 none)
  int32_t* lane = nullptr;     // [ntiles][64 maxw] local cell index of every lane (-1 padding)
  int32_t* wrow = nullptr;     // [ntiles][maxw] parameter row of every wave (-1 idle)
  double* nodeacc = nullptr;   // [ntiles][S][3] per-slot gradient sums
  double* partials = nullptr;  // [kPartialStride][ntiles] energy and virial per tile
  int32_t* sys_ptr = nullptr;  // [nsys + 1] CSR of the tiles of every system
  int32_t* sys_idx = nullptr;  // [ntiles]
  int32_t* hptr = nullptr;     // [N + 1] node -> slot entries
  int32_t* hent = nullptr;     // [E] tile * S + slot, tiles in increasing order
};

// Device-resident force field (include/micmec_amd.h).
struct mm_field {
  int device = 0;
  int model = 0;
  int64_t C = 0, N = 0, W = 0;
  int32_t nrows = 0, max_states = 0;
  int32_t* cnode = nullptr;
  int32_t* cwrap = nullptr;
  int32_t* blk_row = nullptr;
  int32_t* ncell = nullptr;
  int32_t* nstates = nullptr;
  double* tab = nullptr;
  double* kt = nullptr;
  double* gcell = nullptr;
  double* partials = nullptr;
  int* nanflag = nullptr;
  int32_t nsys = 1;                   // independent systems sharing the object
  int32_t* blk_sys = nullptr;         // [W/256] system of every cell block
  int32_t* sys_cblk_ptr = nullptr;    // [nsys + 1] CSR of the cell blocks of every system
  int32_t* sys_cblk_idx = nullptr;    // [W/256]
  std::vector<int64_t> slot_of_cell;  // reference cell index -> work slot
  std::vector<int32_t> blk_row_host;  // [W/256] copy of blk_row
  int split = -1;                     // see ForceArgs::split
  hipStream_t side = nullptr;         // multistate instance of the cell pass runs here
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  mm_tiles tiles;
};

namespace mm {
// Tiled cell pass (mm_tile.hip). launch_tile_cells: energies, virials and per-slot gradient
// sums (grad) of every tile; launch_tile_gather: gpos (+)= the node sums (plain force calls).
hipError_t launch_tile_cells(const mm_field* f, const double* pos, const double* rvecs,
                             int64_t rvecs_stride, bool grad, hipStream_t stream);
hipError_t launch_tile_gather(const mm_field* f, double* gpos, int accumulate, hipStream_t stream);
void free_tiles(mm_tiles& t);
// Recompute mm_field::split from the state counts of the rows (host array).
void update_split(mm_field* f, const int32_t* nstates);
// Fill the cell-pass scheduling fields of a ForceArgs from the field.
struct ForceArgs;
void schedule_args(const mm_field* f, ForceArgs& a);
}  // namespace mm

