/* C ABI of the MI355X micromechanical force field (libmicmec_amd.so).
 *
 * Plain pointers and sizes only. Device pointers are HIP device memory of the device the
 * object was created on; streams are hipStream_t passed as void* (NULL = default stream).
 * All compute calls are asynchronous on the given stream; only mm_field_check, the *_read
 * functions and the mm_cg_* calls given a host output synchronise.
 *
 * Status codes: 0 = ok; MM_ERR_ARG -> ValueError/TypeError in the Python layer;
 * MM_ERR_NAN -> ValueError (reference: micmec/pes/mmff.py:135-148); MM_ERR_HIP and
 * MM_ERR_NOMEM -> RuntimeError / MemoryError. mm_last_error() describes the last failure of
 * the calling thread.
 */
#ifndef MICMEC_AMD_H
#define MICMEC_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MM_OK 0
#define MM_ERR_ARG 1
#define MM_ERR_NAN 2
#define MM_ERR_HIP 3
#define MM_ERR_NOMEM 4

#define MM_MODEL_CORNER8 0  /* micmec/pes/nanocell.py:40-132 */
#define MM_MODEL_ORIGINAL 1 /* micmec/pes/nanocell_original.py:46-132 */

/* NaN flag bits reported by mm_field_check */
#define MM_NAN_ENERGY 1
#define MM_NAN_GPOS 2
#define MM_NAN_VTENS 4

typedef struct mm_field mm_field;

/* Version string of the library build. */
const char* mm_version(void);

/* Description of the last error on the calling thread ("" if none). */
const char* mm_last_error(void);

/* Build the device-resident force field of one system.
 * Replaces ForcePartMechanical.__init__ (micmec/pes/mmff.py:207-245): the parameter parse
 * (218-231) arrives as per-row tables, and the N x N image table of get_mic (258-286) as one
 * byte of wrap bits per cell.
 *   surrounding_nodes [ncells][8], surrounding_cells [nnodes][8] (-1 = no cell): the tables of
 *     micmec/system.py:85-107 in reference order;
 *   wrap_bits [ncells]: bit d set when the cell lies in the last layer of periodic axis d;
 *   cell_row [ncells]: parameter row of each cell (0..nrows-1);
 *   nstates [nrows], h0 [nrows][max_states][3][3] (rows = cell vectors),
 *   c0 [nrows][max_states][3][3][3][3], free_energy [nrows][max_states],
 *   temp_eff [nrows] (kelvin), boltzmann (hartree per kelvin).
 * Host pointers; the tables are copied to device `device`. */
int mm_field_create(mm_field** out, int device, int model, int64_t ncells, int64_t nnodes,
                    const int64_t* surrounding_nodes, const int64_t* surrounding_cells,
                    const uint8_t* wrap_bits, const int32_t* cell_row, int32_t nrows,
                    int32_t max_states, const int32_t* nstates, const double* h0,
                    const double* c0, const double* free_energy, const double* temp_eff,
                    double boltzmann);

/* Energy, node gradient and virial at positions pos [nnodes][3] and domain vectors
 * rvecs [3][3] (device pointers; zeros for a non-periodic system).
 * Replaces ForcePartMechanical._internal_compute (micmec/pes/mmff.py:288-323) with
 * deformation (326-403): energy [1] is required; gpos [nnodes][3] and vtens [3][3] may be
 * NULL. accumulate != 0 adds to the outputs instead of overwriting them (the += of
 * ForcePart.compute, mmff.py:142, 148). */
int mm_field_compute(mm_field* f, const double* pos, const double* rvecs, double* energy,
                     double* gpos, double* vtens, int accumulate, void* stream);

/* mm_field_compute without any allocation, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * vertex-gradient records and block partials go to a static workspace of 64 MiB that every
 * GPU holds for the life of the process, instead of the field's own buffers. No counterpart in
 * the reference. Same arguments, results (bit for bit) and semantics as mm_field_compute;
 * fields needing more than the workspace (about 400,000 cells) are refused with MM_ERR_ARG.
 * Calls that overlap on one GPU (other streams, other threads) share the workspace and race:
 * serialise them. */
int mm_field_compute_static(mm_field* f, const double* pos, const double* rvecs, double* energy,
                            double* gpos, double* vtens, int accumulate, void* stream);

/* Synchronise `stream`, report and clear the NaN flags of the computes issued since the last
 * check (MM_NAN_* bits in *flags). Returns MM_ERR_NAN if any flag was set
 * (micmec/pes/mmff.py:135-148). */
int mm_field_check(mm_field* f, void* stream, int32_t* flags);

/* Batched objects: nsys independent systems in one device object (cell_sys [ncells] gives the
 * system of every cell; node indices are global). No reference counterpart (the reference
 * holds one system per ForcePartMechanical); each system's results equal those of its own
 * mm_field_create object. */
int mm_batch_field_create(mm_field** out, int device, int model, int32_t nsys, int64_t ncells,
                          int64_t nnodes, const int64_t* surrounding_nodes,
                          const int64_t* surrounding_cells, const uint8_t* wrap_bits,
                          const int32_t* cell_row, const int32_t* cell_sys, int32_t nrows,
                          int32_t max_states, const int32_t* nstates, const double* h0,
                          const double* c0, const double* free_energy, const double* temp_eff,
                          double boltzmann);

/* Energies [nsys], gradient [nnodes][3] (or NULL) and virials [nsys][9] (or NULL) of a batched
 * object at positions pos and per-system domain vectors rvecs [nsys][9]; overwrites. */
int mm_batch_compute(mm_field* f, const double* pos, const double* rvecs, double* energies,
                     double* gpos, double* vtens, void* stream);

/* The cell pass of mm_field_compute alone (per-cell energies, gradients and virial partials
 * into the object's own buffers), for per-kernel timing and profiling; no output is written. */
int mm_field_launch_cells(mm_field* f, const double* pos, const double* rvecs, void* stream);

/* One trial point of a conjugate-gradient line search over the node positions (single-system
 * objects). Replaces the work of one CartesianDOF.fun call (micmec/sampling/dof.py:132-153)
 * together with the vector arithmetic around it in the minimiser that CGOptimizer builds
 * (micmec/sampling/opt.py:153-162, molmod.minimizer's ConjugateGradient and NewtonLineSearch)
 * and the measures of CartesianDOF.check_convergence (dof.py:155-193) should the point be
 * accepted:
 *   if set_dir != 0: d <- beta d - g   (the new search direction; beta = 0 restarts along -g);
 *   xa <- x + alpha d;  energy and gradient ga at xa and the domain rvecs;
 *   out[0] = energy, out[1] = sum_n ga_n . d_n (the slope of the line at alpha),
 *   out[2] = sum_n |ga_n|^2, out[3] = sum_n ga_n . g_n (Polak-Ribiere numerator out[2] - out[3]),
 *   out[4] = max_n |ga_n|^2, out[5] = sum_n |xa_n - x_n|^2, out[6] = max_n |xa_n - x_n|^2.
 * x, g, d, xa, ga are [nnodes][3] and out [7], all device pointers; xa and ga must not alias
 * x, g, d or each other. Asynchronous on `stream`, unless host_out (host, [7]) is given: then
 * the call waits for the results and copies them there. Calls on one object share its scratch,
 * so issue them in order on one stream. */
int mm_cg_trial(mm_field* f, const double* x, const double* g, double* d, double beta,
                int set_dir, double alpha, const double* rvecs, double* xa, double* ga,
                double* out, double* host_out, void* stream);

/* The same measures at given points: out[0] = sum_n |g_n|^2, out[1] = sum_n g_n . g_old_n,
 * out[2] = sum_n |x_n - x_old_n|^2, out[3] = max_n |g_n|^2, out[4] = max_n |x_n - x_old_n|^2
 * (CartesianDOF.check_convergence, micmec/sampling/dof.py:165-181, before the square root and
 * the mean). Device pointers, x, x_old, g, g_old [nnodes][3], out [5]; asynchronous on
 * `stream` unless host_out (host, [5]) is given, as for mm_cg_trial. */
int mm_cg_stats(mm_field* f, const double* x, const double* x_old, const double* g,
                const double* g_old, double* out, double* host_out, void* stream);

/* Sizes of the object: cells, nodes, padded work slots. */
int mm_field_info(const mm_field* f, int64_t* ncells, int64_t* nnodes, int64_t* nslots);

void mm_field_destroy(mm_field* f);

/* Replace the cell-type parameters of an existing object (same number of rows and at most
 * max_states states per row), with the elasticity given as 6x6 Voigt matrices:
 * cvoigt [nrows][max_states][6][6] with u = e.Cv.e / 2 for the engineering strain
 * e = (eps00, eps11, eps22, 2 eps12, 2 eps02, 2 eps01); for a tensor C0 of full symmetry
 * Cv[I][J] = C0[i][j][k][l] with (ij) = I and (kl) = J in that order (micmec/analysis/tensor.py
 * voigt map). Other arguments as in mm_field_create. Synchronises `stream` first, then copies
 * the new table; later work on any stream sees it. No reference counterpart: the reference
 * reads parameters once (micmec/pes/mmff.py:218-231); this is the entry point of parameter
 * fitting. */
int mm_field_set_params(mm_field* f, const int32_t* nstates, const double* h0,
                        const double* cvoigt, const double* free_energy, const double* temp_eff,
                        double boltzmann, void* stream);

/* mm_field_set_params with h0, cvoigt, free_energy and temp_eff in device memory (nstates stays
 * a host array): a kernel on `stream` builds the table with the host builder's arithmetic (the
 * same bits), so nothing synchronises unless nstates differs from the object's current counts.
 * Later work must be ordered after `stream`. A singular or non-finite h0 is not reported here:
 * it gives non-finite table entries and then non-finite energies (mm_field_check). No
 * reference counterpart (parameter fitting, as mm_field_set_params). */
int mm_field_set_params_device(mm_field* f, const int32_t* nstates, const double* h0,
                               const double* cvoigt, const double* free_energy,
                               const double* temp_eff, double boltzmann, void* stream);

/* ------------------------------------------------------------------------------------------
 * Differentiation form. No reference counterpart for the parameter derivatives; the position
 * and deformation derivatives agree with ForcePartMechanical (micmec/pes/mmff.py:288-323) and
 * with the automatic derivatives of micmec/pes/mmff_jax.py:249-250.
 *
 * mm_adjoint_create allocates the work buffers of an object (about (2 * npar + 38) / 32 + 24
 * doubles per work slot, and 9 per node) and returns npar = max_states * 32 + 1, the parameter
 * entries per row.
 *
 * mm_adjoint_compute evaluates, at positions pos [nnodes][3] and domain vectors rvecs [nsys][9]:
 *   sys_out [nsys][38]: energy, virial [9], dE/d(domain vectors) [9] (row-major, vector d at
 *     3d..3d+2), then the tangents of those 19 numbers;
 *   gpos [nnodes][3] (or NULL) and, with a direction, its tangent dgpos (or NULL);
 *   params [nrows][npar][2] (value, tangent): per state s of the row, entries s*32 + q:
 *     q = 0..20  dE/dCv: 1/2 sum e_I e_J, (I <= J) row by row of the upper triangle; the
 *                gradient of the 6x6 matrix has this value at (I,J) and at (J,I)
 *     q = 21..29 dE/dP, P = h0^-1 row-major;   q = 30 dE/dw, w = det(h0)/8 (corner8) or
 *                det(h0) (original);          q = 31 dE/d(free energy)
 *   and entry max_states*32: dE/d(kT), kT = boltzmann * temp_eff.
 * The tangent direction (dpos [nnodes][3], drvecs [nsys][9], both or neither) turns on dual
 * numbers: every tangent is the derivative of its value along (dpos, drvecs). energy_weight
 * [nsys] (or NULL = all 1) scales each system's share of the parameter values (not of their
 * tangents). All pointers are device pointers; outputs are overwritten. The vertex gradients
 * are recomputed from the positions on every call: nothing is stored between calls. */
typedef struct mm_adjoint mm_adjoint;
int mm_adjoint_create(mm_adjoint** out, mm_field* field, int64_t* npar);
int mm_adjoint_compute(mm_adjoint* a, const double* pos, const double* dpos, const double* rvecs,
                       const double* drvecs, const double* energy_weight, double* sys_out,
                       double* gpos, double* dgpos, double* params, void* stream);

/* Vector-Jacobian product of (E [nsys], gpos [nnodes][3], vtens [nsys][9]) for the cotangents
 * gE [nsys] (required; zeros for none), gG [nnodes][3] and gV [nsys][9] (NULL = zero): one
 * mm_adjoint_compute along the direction (gG + pos gV^T, rvecs gV^T) with energy_weight gE,
 * framed by two node passes. Outputs: g_pos [nnodes][3] = gE gpos + d(gpos) + gpos gV; sys_out
 * and params as for mm_adjoint_compute (the position-independent rest of the product is
 * small: dL/drvecs = gE dE/drvecs + d(dE/drvecs) + (dE/drvecs) gV and dL/dtheta = the sum of
 * the value and tangent rows of params). node_sys [nnodes] int32 gives each node's system.
 * Uses work buffers of the object (9 nnodes + 9 nsys doubles, made by mm_adjoint_create). No
 * reference counterpart (differentiation form). */
int mm_adjoint_vjp(mm_adjoint* a, const double* pos, const double* rvecs,
                   const int32_t* node_sys, const double* gE, const double* gG, const double* gV,
                   double* sys_out, double* g_pos, double* params, void* stream);

/* Gradients of the parameter tensors from parameter entries as mm_adjoint_compute lays them
 * out: entry q of row r at entries[(r * npar + q) * stride], plus the next double when
 * add_next (stride 2, add_next 1: the value and tangent rows of params summed, the cotangent of
 * a vector-Jacobian product). With h0 [nrows][max_states][9] (row-major, the rest cells the
 * table was built from): g_cv [nrows][max_states][36] (the symmetric 6x6 Voigt gradient),
 * g_h0 [nrows][max_states][9] = -P^T (dE/dP) P^T + (dE/dw) w P^T with P = h0^-1 (cofactors)
 * and w = det(h0)/8 (corner8) or det(h0), g_free [nrows][max_states], all zero for the states
 * a row does not have, and g_teff [nrows] = boltzmann dE/d(kT). Device pointers, one kernel.
 * No reference counterpart (differentiation form). */
int mm_adjoint_param_grads(mm_adjoint* a, const double* entries, int64_t stride, int add_next,
                           const double* h0, double boltzmann, double* g_cv, double* g_h0,
                           double* g_free, double* g_teff, void* stream);
void mm_adjoint_destroy(mm_adjoint* a);

/* ------------------------------------------------------------------------------------------
 * Device MD engine: velocity Verlet with an optional Nose-Hoover chain and MTK barostat,
 * replacing VerletIntegrator.propagate (micmec/sampling/verlet.py:140-190), NHChain /
 * NHCThermostat (micmec/sampling/nvt.py:361-532), TBCombination (micmec/sampling/npt.py:52-177)
 * and MTKBarostat (micmec/sampling/npt.py:513-757). All state lives in device memory owned by
 * the caller: positions, velocities, gradients [N][3], masses [N], a copy of the positions of
 * the previous step [N][3], and a block of MM_S_COUNT doubles (indices below) holding every
 * scalar of the integrator and its hooks.
 * ------------------------------------------------------------------------------------------ */

#define MM_THERMO_NONE 0
#define MM_THERMO_NHC 1
#define MM_THERMO_LANGEVIN 2 /* LangevinThermostat (micmec/sampling/nvt.py:165-218) */
#define MM_BARO_NONE 0
#define MM_BARO_MTK 1
#define MM_BARO_LANGEVIN 2 /* LangevinBarostat (micmec/sampling/npt.py:295-510) */

/* Scalar block layout (doubles). Integer-valued entries are stored as doubles. */
#define MM_S_TIMESTEP 0
#define MM_S_NDOF 1
#define MM_S_BOLTZMANN 2
#define MM_S_THERMO 3 /* MM_THERMO_* of the hook */
#define MM_S_BARO 4   /* MM_BARO_* of the hook */
#define MM_S_CHAIN_LEN 5
#define MM_S_THERMO_TEMP 6
#define MM_S_BARO_TEMP 7
#define MM_S_BARO_PRESS 8
#define MM_S_MASS_PRESS 9
#define MM_S_BARO_NDOF 10
#define MM_S_EXT_PRESS 11
#define MM_S_NVEC 12
#define MM_S_CHAIN_NDOF 13 /* degrees of freedom of the NHC chain (nvt.py:395-401) */
#define MM_S_CHAIN_MASS 16 /* 8 */
#define MM_S_CHAIN_POS 24  /* 8 */
#define MM_S_CHAIN_VEL 32  /* 8 */
#define MM_S_VEL_PRESS 40  /* 9 */
#define MM_S_RVECS 49      /* 9 */
#define MM_S_VOLUME 58
#define MM_S_EPOT 59
#define MM_S_VTENS 60 /* 9 */
#define MM_S_EKIN 69
#define MM_S_SVV 70 /* 9: sum_n m_n v_n (x) v_n */
#define MM_S_TIME 79
#define MM_S_COUNTER 80
#define MM_S_ECONS_THERMO 81
#define MM_S_ECONS_BARO 82
#define MM_S_ECONS_CORR 83
#define MM_S_ETOT 84
#define MM_S_ECONS 85
#define MM_S_TEMP 86
#define MM_S_PRESS 87
#define MM_S_PTENS 88 /* 9 */
#define MM_S_CE 97    /* 5: count, ekin mean, ekin var sum, econs mean, econs var sum */
#define MM_S_CONS_ERR 102
#define MM_S_VMAT 103 /* 9: pending velocity matrix (applied after VSCALE_A, before VSCALE_B) */
#define MM_S_VMAT_ON 112
#define MM_S_XMAT 113 /* 9: pending position transform */
#define MM_S_CHAINVEL0 122
#define MM_S_HAS_CHAINVEL0 123
#define MM_S_RMSD_GPOS 124
#define MM_S_RMSD_DELTA 125
#define MM_S_NAN 126
#define MM_S_EPOT_MECH 127
#define MM_S_VSCALE_A 128 /* pending velocity factor applied first (thermostat post) */
#define MM_S_VSCALE_A_ON 129
#define MM_S_VSCALE_B 130 /* pending velocity factor applied last (thermostat pre) */
#define MM_S_VSCALE_B_ON 131
#define MM_S_NNODES 132 /* nodes of the system (mean of the rmsd values) */
/* Langevin hooks. Their Gaussian noise is Philox4x64-10 with the key (seed, stream) and the
 * counter (node or system, 2 * counter + post, update, block); the seeds are 64-bit words held
 * as two doubles with their 32-bit halves. */
#define MM_S_LT_C1 133     /* exp(-dt / timecon / 2) */
#define MM_S_LT_A 134      /* (1 - c1^2) T k_B; the noise amplitude of node n is sqrt(A / m_n) */
#define MM_S_LT_SEED 135   /* 2 */
#define MM_S_LB_SEED 137   /* 2 */
#define MM_S_LB_TIMECON 139
#define MM_S_LB_RSCALE 140 /* sqrt(2 W k_B T / (dt timecon)), the scale of the random force */
#define MM_S_BARO_ANISO 141   /* 1: 3x3 barostat velocity, 0: isotropic (VEL_PRESS[0]) */
#define MM_S_BARO_VOLCON 142  /* 1: volume-preserving cell fluctuations */
#define MM_S_BARO_EPOT0 143   /* epot and ekin at the start of the barostat call */
#define MM_S_BARO_EKIN0 144
/* The MTK barostat's own Nose-Hoover chain (MTKBarostat(baro_thermo=...), npt.py:601-651),
 * acting on the barostat velocity only; its degrees of freedom are the barostat's. */
#define MM_S_BCHAIN_ON 145
#define MM_S_BCHAIN_LEN 146
#define MM_S_BCHAIN_TEMP 147
#define MM_S_BCHAIN_NDOF 148
/* A second pending velocity matrix, applied after VMAT: the next barostat call's transform
 * while the previous call's still waits (steps of one mm_md_step call). */
#define MM_S_VMAT2 149    /* 9 */
#define MM_S_VMAT2_ON 158
/* A pending position transform applied before XMAT: the barostat post call's scaling of a step
 * that another step of the same call follows, whose positions the next scaling pass brings up
 * to date (the post call's force evaluation applies it on the fly). */
#define MM_S_XMAT2_ON 159
#define MM_S_BCHAIN_MASS 160 /* 8 */
#define MM_S_BCHAIN_POS 168  /* 8 */
#define MM_S_BCHAIN_VEL 176  /* 8 */
#define MM_S_XMAT2 184       /* 9 */
#define MM_S_COUNT 200

/* ------------------------------------------------------------------------------------------
 * Topology of a grid of cells, built on the device: Grid + build_system (micmec/utils.py:55-263)
 * for grids too large for a host pass. grid [nx][ny][nz] (host, x-major) holds the type of every
 * grid point (0 = empty, 1..ntypes); pbc_bits: bit d = periodic along axis d; type_mass
 * [ntypes + 1] the mass of every type (entry 0 unused); spacing [3] the lattice spacing
 * (the diagonal of the majority type's first cell matrix, utils.py:220-233). The tables equal
 * build_system's: cells and nodes in its order, corner k of a cell at offset
 * (000,100,010,001,110,101,011,111), -1 for missing cells, node masses summed in (cell, corner)
 * order, plus the wrap bits of the force kernels and the grid index of every cell.
 * ------------------------------------------------------------------------------------------ */
typedef struct mm_topology mm_topology;

int mm_topology_build(mm_topology** out, int device, const int32_t* grid, int32_t nx,
                      int32_t ny, int32_t nz, int32_t pbc_bits, const double* type_mass,
                      int32_t ntypes, const double* spacing, void* stream);
int mm_topology_sizes(const mm_topology* t, int64_t* ncells, int64_t* nnodes, int64_t* nboundary);
/* Copy the tables out (any pointer may be NULL to skip that table): surrounding_nodes [C][8],
 * surrounding_cells [N][8], boundary_nodes [B], masses [N], pos [N][3], types [C],
 * wrap_bits [C], cell_point [C] (linear grid index). Host or device destinations. */
int mm_topology_copy(const mm_topology* t, int64_t* surrounding_nodes, int64_t* surrounding_cells,
                     int64_t* boundary_nodes, double* masses, double* pos, int32_t* types,
                     uint8_t* wrap_bits, int32_t* cell_point, void* stream);
void mm_topology_destroy(mm_topology* t);

typedef struct mm_md mm_md;

/* Bind an integrator to a force field and to caller-owned device buffers. thermo/baro select
 * the hooks (MM_THERMO_*, MM_BARO_*); their parameters are read from the scalar block at every
 * step. No allocation of the caller's buffers happens here. */
int mm_md_create(mm_md** out, mm_field* field, double* pos, double* vel, double* gpos,
                 const double* masses, double* posold, double* scalars, int thermo, int baro);

/* Batched form: nsys independent systems integrated together. The field comes from
 * mm_batch_field_create; node arrays hold the systems one after another, system s in rows
 * [sys_node_start[s], sys_node_start[s + 1]) with every start a multiple of 256 (rows between
 * a system's last node and the next start are inert padding: mass 1, zero velocity, referenced
 * by no cell). scalars holds nsys blocks of MM_S_COUNT doubles. Every call below then advances
 * or evaluates all systems at once, one launch per kernel. No reference counterpart: the
 * reference integrates one system per process. */
int mm_md_create_batch(mm_md** out, mm_field* field, int32_t nsys, const int64_t* sys_node_start,
                       double* pos, double* vel, double* gpos, const double* masses,
                       double* posold, double* scalars, int thermo, int baro);

/* Force evaluation of VerletIntegrator.initialize (verlet.py:123-125) and of
 * MTKBarostat.init (npt.py:612-614): gpos, epot and, if with_vtens, the virial. */
int mm_md_init_force(mm_md* md, int with_vtens, void* stream);

/* compute_properties of VerletIntegrator.initialize (verlet.py:136, 171-190): kinetic energy,
 * temperature, conserved energy, pressure tensor; counter and time unchanged. */
int mm_md_init_properties(mm_md* md, void* stream);

/* Enqueue nsteps integration steps (verlet.py:140-166 with the hooks' pre/post). With
 * use_graph != 0 one step is captured once into a HIP graph and replayed; the graph is
 * captured again when new state counts (mm_field_set_params) have changed the field's split
 * between multistate and single-state cell blocks. The closing scalar update of each step but
 * the last runs in one kernel with the opening one of the next (barostat and Nose-Hoover chain
 * runs), and with a barostat every step but the last leaves gpos alone after its pre-barostat
 * force evaluation (the gradients of its main and post-barostat evaluations are read by
 * nothing before the next step's pre-barostat evaluation); after the call gpos, like every
 * other array and scalar, holds exactly what nsteps calls of one step would leave. */
int mm_md_step(mm_md* md, int32_t nsteps, int use_graph, void* stream);

/* Apply pending velocity/position transforms so that pos and vel hold the reference state. */
int mm_md_sync(mm_md* md, void* stream);

/* Copy pos into posold (call before the step whose displacement is reported). */
int mm_md_snapshot(mm_md* md, void* stream);

/* rmsd_gpos and rmsd_delta = sqrt(mean((pos - posold)^2)) into the scalar block
 * (verlet.py:172-173). */
int mm_md_rmsd(mm_md* md, void* stream);

void mm_md_destroy(mm_md* md);

#ifdef __cplusplus
}
#endif

#endif /* MICMEC_AMD_H */
