/* C ABI of the MI355X micromechanical force field (libmicmec_amd.so).
 *
 * Plain pointers and sizes only. Device pointers are HIP device memory of the device the
 * object was created on; streams are hipStream_t passed as void* (NULL = default stream).
 * All compute calls are asynchronous on the given stream; only mm_field_check and the
 * *_read functions synchronise.
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

/* Synchronise `stream`, report and clear the NaN flags of the computes issued since the last
 * check (MM_NAN_* bits in *flags). Returns MM_ERR_NAN if any flag was set
 * (micmec/pes/mmff.py:135-148). */
int mm_field_check(mm_field* f, void* stream, int32_t* flags);

/* Sizes of the object: cells, nodes, padded work slots. */
int mm_field_info(const mm_field* f, int64_t* ncells, int64_t* nnodes, int64_t* nslots);

void mm_field_destroy(mm_field* f);

#ifdef __cplusplus
}
#endif

#endif /* MICMEC_AMD_H */
