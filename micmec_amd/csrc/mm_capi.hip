// C ABI: device-resident force field objects (see include/micmec_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/micmec_amd.h"
#include "mm_device.h"
#include "mm_kernels.h"
#include "mm_internal.h"

namespace {
thread_local std::string g_last_error;
}

// Workspace of mm_field_compute_static: the vertex-gradient records (12 planes of the field's
// S slots) followed by the block partials (kPartialStride rows of W / kCellBlock) of one force
// evaluation. Every GPU that loads the library holds its own copy, at its own address, until
// the process ends; in host code the name only stands in for it (hipGetSymbolAddress gives
// the device address, once per device, below). All calls on one GPU share it.
constexpr size_t kStaticScratchDoubles = (size_t)8 << 20;  // 64 MiB
static __device__ __align__(256) double scratch[kStaticScratchDoubles];

namespace mm {

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_error(hipError_t e, const char* what) {
  return set_error(MM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

double inverse3(const double* h, double* inv) { return inverse3_cof(h, inv); }

// Voigt pair (i, j) of index I: 0:(0,0) 1:(1,1) 2:(2,2) 3:(1,2) 4:(0,2) 5:(0,1)
// (micmec/analysis/tensor.py:93).
static const int kVoigt[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};

// Cv[I][J] = mean of C[i][j][k][l] over the orderings (i,j),(j,i) x (k,l),(l,k). With
// engineering strains e this gives eps:C:eps = e.Cv.e and (C:eps)_ij -> (Cv e)_I for symmetric
// eps, whatever symmetry C has.
void voigt_average(const double* C, double* Cv) {
  for (int I = 0; I < 6; ++I) {
    for (int J = 0; J < 6; ++J) {
      const int i = kVoigt[I][0], j = kVoigt[I][1], k = kVoigt[J][0], l = kVoigt[J][1];
      const double s = C[((i * 3 + j) * 3 + k) * 3 + l] + C[((j * 3 + i) * 3 + k) * 3 + l] +
                       C[((i * 3 + j) * 3 + l) * 3 + k] + C[((j * 3 + i) * 3 + l) * 3 + k];
      Cv[6 * I + J] = 0.25 * s;
    }
  }
}

int build_param_table(int model, int32_t nrows, int32_t max_states, const int32_t* nstates,
                      const double* h0, const double* c0, const double* free_energy,
                      const double* temp_eff, double boltzmann, std::vector<double>& tab,
                      std::vector<double>& kt, bool c_is_voigt) {
  tab.assign((size_t)nrows * max_states * kTabStride, 0.0);
  kt.assign(nrows, 0.0);
  for (int t = 0; t < nrows; ++t) {
    if (nstates[t] < 1 || nstates[t] > max_states)
      return set_error(MM_ERR_ARG, "row %d has %d states (1..%d allowed)", t, nstates[t],
                       max_states);
    kt[t] = boltzmann * temp_eff[t];
    if (nstates[t] > 1 && !(kt[t] > 0.0))
      return set_error(MM_ERR_ARG, "row %d: multistate cells need a positive effective_temp", t);
    for (int s = 0; s < nstates[t]; ++s) {
      const size_t ts = (size_t)t * max_states + s;
      double* e = &tab[ts * kTabStride];
      const double det = inverse3(h0 + ts * 9, e);
      if (!(std::fabs(det) > 0.0) || !std::isfinite(det))
        return set_error(MM_ERR_ARG, "row %d state %d: singular equilibrium cell matrix", t, s);
      e[9] = model == MM_MODEL_CORNER8 ? 0.125 * det : det;
      e[10] = free_energy[ts];
      if (c_is_voigt) {
        for (int q = 0; q < 36; ++q) e[kCvOffset + q] = c0[ts * 36 + q];
      } else {
        voigt_average(c0 + ts * 81, e + kCvOffset);
      }
      for (int q = 0; q < 36; ++q) e[kCvwOffset + q] = e[9] * e[kCvOffset + q];
    }
  }
  return MM_OK;
}

void update_split(mm_field* f, const int32_t* nstates) {
  const int64_t nb = (int64_t)f->blk_row_host.size();
  int64_t lead = 0;
  while (lead < nb && nstates[f->blk_row_host[lead]] > 1) ++lead;
  bool ok = true;
  for (int64_t b = lead; b < nb && ok; ++b) ok = nstates[f->blk_row_host[b]] <= 1;
  f->split = ok ? (int)lead : -1;
}

void schedule_args(const mm_field* f, ForceArgs& a) {
  a.split = f->split;
  a.side = f->side;
  a.ev_fork = f->ev_fork;
  a.ev_join = f->ev_join;
}

}  // namespace mm

using namespace mm;


namespace {

template <typename T>
int dmalloc(T** p, size_t count) {
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) return hip_error(e, "hipMalloc");
  return MM_OK;
}

template <typename T>
int h2d(T* dst, const T* src, size_t count) {
  if (count == 0) return MM_OK;
  hipError_t e = hipMemcpy(dst, src, count * sizeof(T), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_error(e, "hipMemcpy");
  return MM_OK;
}

void free_field(mm_field* f) {
  if (!f) return;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(f->device);
  (void)hipFree(f->cnode);
  (void)hipFree(f->cwrap);
  (void)hipFree(f->blk_row);
  (void)hipFree(f->ncell);
  (void)hipFree(f->nrec);
  (void)hipFree(f->cbase);
  (void)hipFree(f->nstates);
  (void)hipFree(f->tab);
  (void)hipFree(f->kt);
  (void)hipFree(f->gcell);
  (void)hipFree(f->partials);
  (void)hipFree(f->cgpart);
  if (f->cghost) (void)hipHostFree(f->cghost);
  (void)hipFree(f->nanflag);
  (void)hipFree(f->blk_sys);
  (void)hipFree(f->sys_cblk_ptr);
  (void)hipFree(f->sys_cblk_idx);
  if (f->side) {
    (void)hipStreamSynchronize(f->side);
    (void)hipStreamDestroy(f->side);
  }
  if (f->ev_fork) (void)hipEventDestroy(f->ev_fork);
  if (f->ev_join) (void)hipEventDestroy(f->ev_join);
  (void)hipSetDevice(cur);
  delete f;
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    (void)hipGetDevice(&prev);
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

// The topology, tables and scratch of a field object in a ForceArgs; positions, domain and
// outputs are left to the caller.
void field_args(const mm_field* f, ForceArgs& a) {
  a.cnode = f->cnode;
  a.cwrap = f->cwrap;
  a.blk_row = f->blk_row;
  a.tab = f->tab;
  a.kt = f->kt;
  a.nstates = f->nstates;
  a.max_states = f->max_states;
  a.W = f->W;
  a.N = f->N;
  a.ncell = f->nrec;
  a.gcell = f->gcell;
  a.rstride = f->S;
  a.rshift = 2;
  a.cbase = f->cbase;
  a.partials = f->partials;
  a.model = f->model;
  a.nanflag = f->nanflag;
  schedule_args(f, a);
}

}  // namespace

extern "C" {

const char* mm_version(void) { return "micmec_amd 0.1 (gfx950, fp64)"; }

const char* mm_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"

namespace {

// Shared construction of single-system and batched objects. cell_sys == nullptr means one system.
// Corner k lies on the upper (+z) face of its cell; its face class (the lower-face corner
// 0, 1, 2, 4 it coincides with in the +z neighbour, or itself) numbers the gradient planes.
constexpr int kOffZ[8] = {0, 0, 0, 1, 0, 1, 1, 1};
constexpr int kFaceClass[8] = {0, 1, 2, 0, 3, 1, 2, 3};

int create_field(mm_field** out, int device, int model, int32_t nsys, int64_t ncells,
                 int64_t nnodes, const int64_t* surrounding_nodes,
                 const int64_t* surrounding_cells, const uint8_t* wrap_bits,
                 const int32_t* cell_row, const int32_t* cell_sys, int32_t nrows,
                 int32_t max_states, const int32_t* nstates, const double* h0, const double* c0,
                 const double* free_energy, const double* temp_eff, double boltzmann) {
  g_last_error.clear();
  if (!out) return set_error(MM_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (model != MM_MODEL_CORNER8 && model != MM_MODEL_ORIGINAL)
    return set_error(MM_ERR_ARG, "unknown model %d", model);
  if (ncells < 0 || nnodes < 0) return set_error(MM_ERR_ARG, "negative sizes");
  if (ncells > (1LL << 27)) return set_error(MM_ERR_ARG, "too many cells (%lld)", (long long)ncells);
  if (nnodes > (1LL << 30)) return set_error(MM_ERR_ARG, "too many nodes");
  if (nsys < 1 || nsys > (1 << 24)) return set_error(MM_ERR_ARG, "bad system count %d", nsys);
  if (nrows < 1 || nrows > 65535 || max_states < 1 || max_states > kMaxStates)
    return set_error(MM_ERR_ARG, "bad parameter table shape %d x %d", nrows, max_states);
  if (ncells > 0 && (!surrounding_nodes || !wrap_bits || !cell_row))
    return set_error(MM_ERR_ARG, "missing cell tables");
  if (nnodes > 0 && !surrounding_cells) return set_error(MM_ERR_ARG, "missing node table");
  // Validate every index the kernels dereference.
  for (int64_t c = 0; c < ncells; ++c) {
    for (int k = 0; k < 8; ++k) {
      const int64_t n = surrounding_nodes[c * 8 + k];
      if (n < 0 || n >= nnodes)
        return set_error(MM_ERR_ARG, "surrounding_nodes[%lld][%d] = %lld out of range",
                         (long long)c, k, (long long)n);
    }
    if (cell_row[c] < 0 || cell_row[c] >= nrows)
      return set_error(MM_ERR_ARG, "cell %lld has parameter row %d", (long long)c, cell_row[c]);
    if (wrap_bits[c] > 7) return set_error(MM_ERR_ARG, "wrap bits of cell %lld > 7", (long long)c);
    if (cell_sys && (cell_sys[c] < 0 || cell_sys[c] >= nsys))
      return set_error(MM_ERR_ARG, "cell %lld belongs to system %d", (long long)c, cell_sys[c]);
  }
  for (int64_t n = 0; n < nnodes; ++n) {
    for (int k = 0; k < 8; ++k) {
      const int64_t c = surrounding_cells[n * 8 + k];
      if (c < -1 || c >= ncells)
        return set_error(MM_ERR_ARG, "surrounding_cells[%lld][%d] = %lld out of range",
                         (long long)n, k, (long long)c);
    }
  }
  // The two tables must describe the same incidence: node n is corner k of cell c exactly when
  // surrounding_cells[n][k] = c (the cell pass folds the shared faces of neighbouring cells
  // through surrounding_nodes, the gather finds the sums through surrounding_cells).
  int64_t listed = 0;
  for (int64_t n = 0; n < nnodes; ++n) {
    for (int k = 0; k < 8; ++k) {
      const int64_t c = surrounding_cells[n * 8 + k];
      if (c < 0) continue;
      ++listed;
      if (surrounding_nodes[c * 8 + k] != n)
        return set_error(MM_ERR_ARG, "surrounding_cells[%lld][%d] = %lld but corner %d of that "
                         "cell is node %lld", (long long)n, k, (long long)c, k,
                         (long long)surrounding_nodes[c * 8 + k]);
    }
  }
  if (listed != 8 * ncells)
    return set_error(MM_ERR_ARG, "surrounding_cells lists %lld cell corners, the cells have %lld",
                     (long long)listed, (long long)(8 * ncells));
  std::vector<double> tab, kt;
  int rc = build_param_table(model, nrows, max_states, nstates, h0, c0, free_energy, temp_eff,
                             boltzmann, tab, kt, false);
  if (rc) return rc;

  // Work order: cells grouped by (parameter row, system), reference order inside a group, each
  // group padded to a whole number of blocks so that every block is type- and system-uniform.
  // Rows with several states come first: their blocks take longest, and starting them early
  // lets the single-state blocks fill the end of each launch.
  std::vector<std::vector<int64_t>> groups((size_t)nrows * nsys);
  for (int64_t c = 0; c < ncells; ++c)
    groups[(size_t)cell_row[c] * nsys + (cell_sys ? cell_sys[c] : 0)].push_back(c);
  int64_t W = 0;
  for (const auto& g : groups) W += (int64_t)((g.size() + kCellBlock - 1) / kCellBlock) * kCellBlock;
  // Records are 32-bit: 8 * slot + corner in the per-corner layout, 4 * (W + i) + class in the
  // face-class layout (i < W), so 8 * W must fit. Padding every (row, system) group to a
  // whole block can reach that with many small systems long before the cell count limit.
  if (8 * W > (int64_t)INT32_MAX)
    return set_error(MM_ERR_ARG, "%lld work slots after padding %lld (row, system) groups to "
                     "blocks of %d exceed the 32-bit record range; split the batch",
                     (long long)W, (long long)groups.size(), kCellBlock);
  std::vector<int32_t> cnode((size_t)8 * W, -1), cwrap(W, 0), blk_row(W / kCellBlock, 0),
      blk_sys(W / kCellBlock, 0);
  std::vector<int64_t> slot(ncells, -1);
  std::vector<std::vector<int32_t>> sys_blocks(nsys);
  std::vector<int> row_order;
  for (int t = 0; t < nrows; ++t)
    if (nstates[t] > 1) row_order.push_back(t);
  for (int t = 0; t < nrows; ++t)
    if (nstates[t] <= 1) row_order.push_back(t);
  int64_t w0 = 0;
  for (int t : row_order) {
    for (int sy = 0; sy < nsys; ++sy) {
      const auto& g = groups[(size_t)t * nsys + sy];
      const int64_t nb = (int64_t)((g.size() + kCellBlock - 1) / kCellBlock);
      for (int64_t b = 0; b < nb; ++b) {
        blk_row[w0 / kCellBlock + b] = t;
        blk_sys[w0 / kCellBlock + b] = sy;
        sys_blocks[sy].push_back((int32_t)(w0 / kCellBlock + b));
      }
      for (size_t i = 0; i < g.size(); ++i) {
        const int64_t c = g[i];
        const int64_t w = w0 + (int64_t)i;
        slot[c] = w;
        for (int k = 0; k < 8; ++k) cnode[(size_t)w * 8 + k] = (int32_t)surrounding_nodes[c * 8 + k];
        cwrap[w] = wrap_bits[c];
      }
      w0 += nb * kCellBlock;
    }
  }
  std::vector<int32_t> sys_ptr(nsys + 1, 0), sys_idx;
  for (int sy = 0; sy < nsys; ++sy) {
    sys_ptr[sy + 1] = sys_ptr[sy] + (int32_t)sys_blocks[sy].size();
    sys_idx.insert(sys_idx.end(), sys_blocks[sy].begin(), sys_blocks[sy].end());
  }
  std::vector<int32_t> ncell((size_t)8 * nnodes, -1);
  for (int64_t n = 0; n < nnodes; ++n) {
    for (int k = 0; k < 8; ++k) {
      const int64_t c = surrounding_cells[n * 8 + k];
      if (c >= 0) ncell[(size_t)k * nnodes + n] = (int32_t)(slot[c] * 8 + k);
    }
  }
  // Shared faces along z, the axis along which the reference numbers consecutive cells: when
  // slot w + 1 of the same wave holds the +z neighbour of slot w (corner 3 of w is corner 0 of
  // w + 1), the cell pass adds w's upper-face vertex gradients (corners 3, 5, 6, 7) to w + 1's
  // lower-face ones (corners 0, 1, 2, 4) with a lane shuffle (cwrap bit 4 on w, bit 3 on
  // w + 1). Lower faces are stored per slot in four class planes; the upper faces that are not
  // folded go to compact slots W + i after the work slots, numbered per wave in lane order
  // from cbase. nrec lists every node's records (slot << 2 | class), skipping folded faces.
  std::vector<int64_t> cell_of(W, -1);
  for (int64_t c = 0; c < ncells; ++c) cell_of[slot[c]] = c;
  // The fold needs the whole face shared: upper corners 3, 5, 6, 7 of w are the nodes of
  // lower corners 0, 1, 2, 4 of w + 1, pair by pair (the face class numbering).
  constexpr int kUpper[4] = {3, 5, 6, 7}, kLower[4] = {0, 1, 2, 4};
  for (int64_t w = 0; w + 1 < W; ++w) {
    const int64_t c = cell_of[w], up = cell_of[w + 1];
    if (c < 0 || up < 0 || (w + 1) % 64 == 0) continue;
    bool shared = true;
    for (int q = 0; q < 4; ++q)
      shared = shared && surrounding_nodes[c * 8 + kUpper[q]] == surrounding_nodes[up * 8 + kLower[q]];
    if (shared) {
      cwrap[w] |= 16;
      cwrap[w + 1] |= 8;
    }
  }
  std::vector<int32_t> cbase(W / 64, 0);
  std::vector<int64_t> cidx(W, -1);
  int64_t ncomp = 0;
  for (int64_t v = 0; v < W / 64; ++v) {
    cbase[v] = (int32_t)ncomp;
    for (int64_t w = 64 * v; w < 64 * v + 64; ++w)
      if (cell_of[w] >= 0 && !(cwrap[w] & 16)) cidx[w] = ncomp++;
  }
  const int64_t S = W + ncomp;
  std::vector<int32_t> nrec((size_t)8 * nnodes, -1);
  for (int64_t n = 0; n < nnodes; ++n) {
    for (int k = 0; k < 8; ++k) {
      const int64_t c = surrounding_cells[n * 8 + k];
      if (c < 0) continue;
      const int64_t w = slot[c];
      const int upper = kOffZ[k], cls = kFaceClass[k];
      if (!upper) nrec[(size_t)k * nnodes + n] = (int32_t)(4 * w + cls);
      else if (!(cwrap[w] & 16)) nrec[(size_t)k * nnodes + n] = (int32_t)(4 * (W + cidx[w]) + cls);
    }
  }

  DeviceGuard guard(device);
  mm_field* f = new mm_field();
  f->device = device;
  f->model = model;
  f->C = ncells;
  f->N = nnodes;
  f->W = W;
  f->S = S;
  f->nrows = nrows;
  f->max_states = max_states;
  f->nsys = nsys;
  f->slot_of_cell = std::move(slot);
  const int64_t nb = W / kCellBlock;
  if ((rc = dmalloc(&f->cnode, (size_t)8 * W)) || (rc = dmalloc(&f->cwrap, W)) ||
      (rc = dmalloc(&f->blk_row, nb)) || (rc = dmalloc(&f->ncell, (size_t)8 * nnodes)) ||
      (rc = dmalloc(&f->nrec, (size_t)8 * nnodes)) || (rc = dmalloc(&f->cbase, W / 64)) ||
      (rc = dmalloc(&f->nstates, nrows)) || (rc = dmalloc(&f->tab, tab.size())) ||
      (rc = dmalloc(&f->kt, kt.size())) || (rc = dmalloc(&f->gcell, (size_t)24 * W)) ||
      (rc = dmalloc(&f->partials, (size_t)kPartialStride * (nb > 0 ? nb : 1))) ||
      (rc = dmalloc(&f->cgpart, (size_t)8 * (nnodes > 0 ? (nnodes + 255) / 256 : 1))) ||
      (rc = dmalloc(&f->nanflag, 1)) || (rc = dmalloc(&f->blk_sys, nb)) ||
      (rc = dmalloc(&f->sys_cblk_ptr, (size_t)nsys + 1)) ||
      (rc = dmalloc(&f->sys_cblk_idx, nb))) {
    free_field(f);
    return rc;
  }
  if ((rc = h2d(f->cnode, cnode.data(), cnode.size())) || (rc = h2d(f->cwrap, cwrap.data(), W)) ||
      (rc = h2d(f->blk_row, blk_row.data(), blk_row.size())) ||
      (rc = h2d(f->ncell, ncell.data(), ncell.size())) ||
      (rc = h2d(f->nrec, nrec.data(), nrec.size())) ||
      (rc = h2d(f->cbase, cbase.data(), cbase.size())) ||
      (rc = h2d(f->nstates, nstates, (size_t)nrows)) ||
      (rc = h2d(f->tab, tab.data(), tab.size())) || (rc = h2d(f->kt, kt.data(), kt.size())) ||
      (rc = h2d(f->blk_sys, blk_sys.data(), blk_sys.size())) ||
      (rc = h2d(f->sys_cblk_ptr, sys_ptr.data(), sys_ptr.size())) ||
      (rc = h2d(f->sys_cblk_idx, sys_idx.data(), sys_idx.size()))) {
    free_field(f);
    return rc;
  }
  hipError_t e = hipMemset(f->nanflag, 0, sizeof(int));
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&f->cghost, 8 * sizeof(double), hipHostMallocDefault);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&f->side, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&f->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&f->ev_join, hipEventDisableTiming);
  if (e != hipSuccess) {
    free_field(f);
    return hip_error(e, "mm_field_create");
  }
  f->blk_row_host = blk_row;
  f->nstates_host.assign(nstates, nstates + nrows);
  update_split(f, nstates);
  *out = f;
  return MM_OK;
}

}  // namespace

extern "C" {

int mm_field_create(mm_field** out, int device, int model, int64_t ncells, int64_t nnodes,
                    const int64_t* surrounding_nodes, const int64_t* surrounding_cells,
                    const uint8_t* wrap_bits, const int32_t* cell_row, int32_t nrows,
                    int32_t max_states, const int32_t* nstates, const double* h0,
                    const double* c0, const double* free_energy, const double* temp_eff,
                    double boltzmann) {
  return create_field(out, device, model, 1, ncells, nnodes, surrounding_nodes,
                      surrounding_cells, wrap_bits, cell_row, nullptr, nrows, max_states,
                      nstates, h0, c0, free_energy, temp_eff, boltzmann);
}

int mm_batch_field_create(mm_field** out, int device, int model, int32_t nsys, int64_t ncells,
                          int64_t nnodes, const int64_t* surrounding_nodes,
                          const int64_t* surrounding_cells, const uint8_t* wrap_bits,
                          const int32_t* cell_row, const int32_t* cell_sys, int32_t nrows,
                          int32_t max_states, const int32_t* nstates, const double* h0,
                          const double* c0, const double* free_energy, const double* temp_eff,
                          double boltzmann) {
  if (ncells > 0 && !cell_sys) return set_error(MM_ERR_ARG, "cell_sys is NULL");
  return create_field(out, device, model, nsys, ncells, nnodes, surrounding_nodes,
                      surrounding_cells, wrap_bits, cell_row, cell_sys, nrows, max_states,
                      nstates, h0, c0, free_energy, temp_eff, boltzmann);
}

int mm_field_set_params(mm_field* f, const int32_t* nstates, const double* h0,
                        const double* cvoigt, const double* free_energy, const double* temp_eff,
                        double boltzmann, void* stream) {
  if (!f || !nstates || !h0 || !cvoigt || !free_energy || !temp_eff)
    return set_error(MM_ERR_ARG, "missing argument");
  std::vector<double> tab, kt;
  int rc = build_param_table(f->model, f->nrows, f->max_states, nstates, h0, cvoigt, free_energy,
                             temp_eff, boltzmann, tab, kt, true);
  if (rc) return rc;
  DeviceGuard guard(f->device);
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  if (e == hipSuccess)
    e = hipMemcpy(f->tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(f->kt, kt.data(), sizeof(double) * kt.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(f->nstates, nstates, sizeof(int32_t) * f->nrows, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_error(e, "mm_field_set_params");
  f->nstates_host.assign(nstates, nstates + f->nrows);
  update_split(f, nstates);
  return MM_OK;
}

int mm_field_set_params_device(mm_field* f, const int32_t* nstates, const double* h0,
                               const double* cvoigt, const double* free_energy,
                               const double* temp_eff, double boltzmann, void* stream) {
  if (!f || !nstates || !h0 || !cvoigt || !free_energy || !temp_eff)
    return set_error(MM_ERR_ARG, "missing argument");
  for (int t = 0; t < f->nrows; ++t)
    if (nstates[t] < 1 || nstates[t] > f->max_states)
      return set_error(MM_ERR_ARG, "row %d has %d states (1..%d allowed)", t, nstates[t],
                       f->max_states);
  DeviceGuard guard(f->device);
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSuccess;
  if (!std::equal(f->nstates_host.begin(), f->nstates_host.end(), nstates)) {
    // a new state count per row: the work order's split changes too (rare; synchronises)
    e = hipStreamSynchronize(s);
    if (e == hipSuccess)
      e = hipMemcpy(f->nstates, nstates, sizeof(int32_t) * f->nrows, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_error(e, "mm_field_set_params_device");
    f->nstates_host.assign(nstates, nstates + f->nrows);
    update_split(f, nstates);
  }
  e = launch_build_table(f->model == MM_MODEL_CORNER8, f->nrows, f->max_states, f->nstates, h0,
                         cvoigt, free_energy, temp_eff, boltzmann, f->tab, f->kt, s);
  if (e != hipSuccess) return hip_error(e, "parameter table kernel");
  return MM_OK;
}

int mm_batch_compute(mm_field* f, const double* pos, const double* rvecs, double* energies,
                     double* gpos, double* vtens, void* stream) {
  if (!f) return set_error(MM_ERR_ARG, "field is NULL");
  if ((f->N > 0 && !pos) || !rvecs || !energies)
    return set_error(MM_ERR_ARG, "pos, rvecs and energies are required");
  DeviceGuard guard(f->device);
  ForceArgs a;
  a.pos = pos;
  a.rvecs = rvecs;
  a.rvecs_stride = 9;
  a.blk_sys = f->blk_sys;
  field_args(f, a);
  a.energy = nullptr;
  a.gpos = gpos;
  a.vtens = vtens;
  a.accumulate = 0;
  a.force_grad = 1;
  hipError_t e = launch_cells(a, (hipStream_t)stream);
  if (e == hipSuccess && gpos) e = launch_gather(a, (hipStream_t)stream);
  if (e == hipSuccess)
    e = launch_batch_finalize(f->partials, (int)(f->W / kCellBlock), f->sys_cblk_ptr,
                              f->sys_cblk_idx, f->nsys, energies, vtens, f->nanflag,
                              (hipStream_t)stream);
  if (e != hipSuccess) return hip_error(e, "batched force kernels");
  return MM_OK;
}

int mm_field_compute(mm_field* f, const double* pos, const double* rvecs, double* energy,
                     double* gpos, double* vtens, int accumulate, void* stream) {
  if (!f) return set_error(MM_ERR_ARG, "field is NULL");
  if (!energy) return set_error(MM_ERR_ARG, "energy output is required");
  if (f->N > 0 && !pos) return set_error(MM_ERR_ARG, "pos is NULL");
  if (!rvecs) return set_error(MM_ERR_ARG, "rvecs is NULL");
  DeviceGuard guard(f->device);
  ForceArgs a;
  a.pos = pos;
  a.rvecs = rvecs;
  field_args(f, a);
  a.energy = energy;
  a.gpos = gpos;
  a.vtens = vtens;
  a.accumulate = accumulate;
  hipError_t e = launch_force(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_error(e, "force kernels");
  return MM_OK;
}

}  // extern "C"

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kScratchDevices] = {};

// Device address of the static workspace on `device` (which must be current), looked up on the
// first call there and kept.
int static_scratch(int device, double** out) {
  if (device < 0 || device >= kScratchDevices)
    return mm::set_error(MM_ERR_ARG, "device %d out of range for the static workspace", device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return mm::hip_error(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[device];
  return MM_OK;
}

}  // namespace

extern "C" {

int mm_field_compute_static(mm_field* f, const double* pos, const double* rvecs, double* energy,
                            double* gpos, double* vtens, int accumulate, void* stream) {
  if (!f) return set_error(MM_ERR_ARG, "field is NULL");
  if (!energy) return set_error(MM_ERR_ARG, "energy output is required");
  if (f->N > 0 && !pos) return set_error(MM_ERR_ARG, "pos is NULL");
  if (!rvecs) return set_error(MM_ERR_ARG, "rvecs is NULL");
  const int64_t nb = f->W / kCellBlock;
  const int64_t need = 12 * f->S + (int64_t)kPartialStride * nb;
  if (need > (int64_t)kStaticScratchDoubles)
    return set_error(MM_ERR_ARG, "this field needs %lld doubles of workspace (%lld cells), the "
                     "static workspace holds %lld; use mm_field_compute",
                     (long long)need, (long long)f->C, (long long)kStaticScratchDoubles);
  DeviceGuard guard(f->device);
  double* ws = nullptr;
  int rc = static_scratch(f->device, &ws);
  if (rc) return rc;
  ForceArgs a;
  a.pos = pos;
  a.rvecs = rvecs;
  field_args(f, a);
  a.gcell = ws;
  a.partials = ws + 12 * f->S;
  a.energy = energy;
  a.gpos = gpos;
  a.vtens = vtens;
  a.accumulate = accumulate;
  hipError_t e = launch_force(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_error(e, "force kernels");
  return MM_OK;
}

int mm_field_check(mm_field* f, void* stream, int32_t* flags) {
  if (!f) return set_error(MM_ERR_ARG, "field is NULL");
  DeviceGuard guard(f->device);
  int h = 0;
  hipError_t e = hipMemcpyAsync(&h, f->nanflag, sizeof(int), hipMemcpyDeviceToHost,
                                (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) return hip_error(e, "reading the NaN flag");
  if (flags) *flags = h;
  if (h) {
    e = hipMemsetAsync(f->nanflag, 0, sizeof(int), (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return hip_error(e, "clearing the NaN flag");
    return set_error(MM_ERR_NAN, "not-a-number in %s%s%s", (h & MM_NAN_ENERGY) ? "energy " : "",
                     (h & MM_NAN_GPOS) ? "gpos " : "", (h & MM_NAN_VTENS) ? "vtens" : "");
  }
  return MM_OK;
}

int mm_field_launch_cells(mm_field* f, const double* pos, const double* rvecs, void* stream) {
  if (!f) return set_error(MM_ERR_ARG, "field is NULL");
  if ((f->N > 0 && !pos) || !rvecs) return set_error(MM_ERR_ARG, "pos or rvecs is NULL");
  DeviceGuard guard(f->device);
  ForceArgs a;
  a.pos = pos;
  a.rvecs = rvecs;
  field_args(f, a);
  a.energy = nullptr;
  a.gpos = nullptr;
  a.vtens = nullptr;
  a.accumulate = 0;
  a.force_grad = 1;
  hipError_t e = launch_cells(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_error(e, "cell kernel");
  return MM_OK;
}

// Copy n results from the device to host_out through the object's pinned staging, after the
// work queued on the stream (a synchronising read).
static int read_results(mm_field* f, const double* out, double* host_out, int n, hipStream_t s) {
  hipError_t e = hipMemcpyAsync(f->cghost, out, sizeof(double) * n, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_error(e, "reading line-search results");
  std::memcpy(host_out, f->cghost, sizeof(double) * n);
  return MM_OK;
}

int mm_cg_trial(mm_field* f, const double* x, const double* g, double* d, double beta,
                int set_dir, double alpha, const double* rvecs, double* xa, double* ga,
                double* out, double* host_out, void* stream) {
  if (!f) return set_error(MM_ERR_ARG, "field is NULL");
  if (f->nsys != 1) return set_error(MM_ERR_ARG, "mm_cg_trial needs a single-system object");
  if (!rvecs || !out) return set_error(MM_ERR_ARG, "rvecs and out are required");
  if (f->N > 0 && (!x || !g || !d || !xa || !ga))
    return set_error(MM_ERR_ARG, "x, g, d, xa and ga are required");
  if (f->N > 0 && (xa == x || xa == g || xa == d || ga == x || ga == g || ga == d || xa == ga))
    return set_error(MM_ERR_ARG, "xa and ga must not alias x, g, d or each other");
  DeviceGuard guard(f->device);
  hipStream_t s = (hipStream_t)stream;
  ForceArgs a;
  a.pos = xa;
  a.rvecs = rvecs;
  field_args(f, a);
  a.energy = out;
  a.gpos = ga;
  a.vtens = nullptr;
  a.accumulate = 0;
  const int64_t nbn = f->N > 0 ? (f->N + 255) / 256 : 0;
  double* tpart = f->cgpart;            // rows 0-1: displacement sum and maximum
  double* gpart = f->cgpart + 2 * nbn;  // rows 2-5: g.d, g.g, g.g_old, max |g_n|^2
  hipError_t e = launch_cg_trial(x, g, d, beta, set_dir, alpha, xa, f->N, tpart, s);
  if (e == hipSuccess) e = launch_cells(a, s);
  if (e == hipSuccess) e = launch_gather_dot(a, d, g, gpart, s);
  if (e == hipSuccess) {
    ReduceList list;
    add_force_totals(list, a, (int)(f->W / kCellBlock));
    const int maxop = nbn > 0 ? kReduceMax : kReduceSum;  // an empty maximum is 0
    for (int q = 0; q < 4; ++q)
      list.add(gpart + q * nbn, nbn, q < 3 ? kReduceSum : maxop, out + 1 + q, 0, MM_NAN_GPOS);
    list.add(tpart, nbn, kReduceSum, out + 5, 0, 0);
    list.add(tpart + nbn, nbn, maxop, out + 6, 0, 0);
    e = launch_reduce(list, f->nanflag, s);
  }
  if (e != hipSuccess) return hip_error(e, "line-search trial kernels");
  return host_out ? read_results(f, out, host_out, 7, s) : MM_OK;
}

int mm_cg_stats(mm_field* f, const double* x, const double* x_old, const double* g,
                const double* g_old, double* out, double* host_out, void* stream) {
  if (!f) return set_error(MM_ERR_ARG, "field is NULL");
  if (!out) return set_error(MM_ERR_ARG, "out is required");
  if (f->N > 0 && (!x || !x_old || !g || !g_old))
    return set_error(MM_ERR_ARG, "x, x_old, g and g_old are required");
  DeviceGuard guard(f->device);
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  if (f->N > 0) {
    e = launch_cg_stats(x, x_old, g, g_old, f->N, f->cgpart, out, f->nanflag, s);
  } else {
    ReduceList list;  // empty sums: five zeros
    for (int q = 0; q < 5; ++q) list.add(f->cgpart, 0, kReduceSum, out + q, 0, 0);
    e = launch_reduce(list, f->nanflag, s);
  }
  if (e != hipSuccess) return hip_error(e, "line-search statistics kernels");
  return host_out ? read_results(f, out, host_out, 5, s) : MM_OK;
}

int mm_field_info(const mm_field* f, int64_t* ncells, int64_t* nnodes, int64_t* nslots) {
  if (!f) return set_error(MM_ERR_ARG, "field is NULL");
  if (ncells) *ncells = f->C;
  if (nnodes) *nnodes = f->N;
  if (nslots) *nslots = f->W;
  return MM_OK;
}

void mm_field_destroy(mm_field* f) { free_field(f); }

}  // extern "C"
