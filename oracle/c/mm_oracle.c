/* CPU restatement of MicMec's force field in C with OpenMP (test infrastructure only).
 *
 * Checker, not product: only tests, the smoke check and the CPU-baseline leg of the benchmark
 * load it. It restates, cell by cell and with the full 3x3x3x3 elasticity tensor:
 *   micmec/pes/mmff.py:326-403        deformation: vertices with image shifts, per-state energy
 *                                     plus free energy, log-sum-exp over states, weighted gradient
 *   micmec/pes/nanocell.py:40-132     eight-corner energy and its gradient through dM_a/dr_v
 *   micmec/pes/nanocell_original.py:38-132  averaged-matrix energy and gradient (1/4 dM_v/dr_v)
 *   micmec/pes/nanocell_utils.py:32-75      corner matrices: row d of M_a is the +d edge at a
 *   micmec/pes/mmff.py:299-323        epot = sum of cell energies, gpos[n] = sum over k = 0..7 of
 *                                     the corner-k gradient of cell surrounding_cells[n][k],
 *                                     vtens = sum over cells and vertices of g (x) r
 * Image shifts follow mmff.py:258-286 through one wrap bit per axis and cell: corner k of a cell
 * in the last layer along axis d with offset_d(k) = 1 is shifted by rvecs[d].
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* corner offsets, micmec/utils.py:43-52 */
static const int kOff[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                               {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};

static int corner_of(int x, int y, int z) {
  for (int k = 0; k < 8; ++k)
    if (kOff[k][0] == x && kOff[k][1] == y && kOff[k][2] == z) return k;
  return -1;
}

/* mult[a][d][v]: row d of M_a = sum_v mult[a][d][v] r_v (the edge along d through corner a,
 * upper end minus lower end) */
static int g_mult[8][3][8];

static void init_mult(void) {
  memset(g_mult, 0, sizeof(g_mult));
  for (int a = 0; a < 8; ++a)
    for (int d = 0; d < 3; ++d) {
      int lo[3] = {kOff[a][0], kOff[a][1], kOff[a][2]};
      int hi[3] = {kOff[a][0], kOff[a][1], kOff[a][2]};
      lo[d] = 0;
      hi[d] = 1;
      g_mult[a][d][corner_of(lo[0], lo[1], lo[2])] -= 1;
      g_mult[a][d][corner_of(hi[0], hi[1], hi[2])] += 1;
    }
}

static double inv3(const double* h, double* out) {
  const double det = h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
                     h[2] * (h[3] * h[7] - h[4] * h[6]);
  out[0] = (h[4] * h[8] - h[5] * h[7]) / det;
  out[1] = (h[2] * h[7] - h[1] * h[8]) / det;
  out[2] = (h[1] * h[5] - h[2] * h[4]) / det;
  out[3] = (h[5] * h[6] - h[3] * h[8]) / det;
  out[4] = (h[0] * h[8] - h[2] * h[6]) / det;
  out[5] = (h[2] * h[3] - h[0] * h[5]) / det;
  out[6] = (h[3] * h[7] - h[4] * h[6]) / det;
  out[7] = (h[1] * h[6] - h[0] * h[7]) / det;
  out[8] = (h[0] * h[4] - h[1] * h[3]) / det;
  return det;
}

/* Green-Lagrange strain of the matrix M (rows = edges): X = (h0inv M)^T,
 * eps = (X^T X - I) / 2 (nanocell.py:64-71); returns X too. */
static void strain_of(const double M[3][3], const double* hi, double X[3][3], double eps[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) {
      double s = 0.0;
      for (int j = 0; j < 3; ++j) s += M[j][i] * hi[k * 3 + j];
      X[i][k] = s;
    }
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) {
      double s = 0.0;
      for (int j = 0; j < 3; ++j) s += X[j][i] * X[j][k];
      eps[i][k] = 0.5 * (s - (i == k ? 1.0 : 0.0));
    }
}

/* sigma = C0 : eps, u = eps : C0 : eps / 2 */
static double stress_of(const double* C, const double eps[3][3], double sig[3][3]) {
  double u = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) s += C[((i * 3 + j) * 3 + k) * 3 + l] * eps[k][l];
      sig[i][j] = s;
      u += 0.5 * eps[i][j] * s;
    }
  return u;
}

/* Contribution of one strained matrix to dE/dr_v[axis]: with m = X[axis][:] (x) (h0inv mu),
 * mu_m = the coefficient of r_v in row m of M, and eps_d = (m + m^T) / 2, it is
 * sum_ij eps_d[j][i] sigma[i][j] (nanocell.py:108-130 with dM/dr written out). */
static double dstrain_dot(const double X[3][3], const double* hi, const double mu[3], int axis,
                          const double sig[3][3]) {
  double hm[3];
  for (int j = 0; j < 3; ++j) hm[j] = hi[j * 3 + 0] * mu[0] + hi[j * 3 + 1] * mu[1] + hi[j * 3 + 2] * mu[2];
  double s = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const double mij = X[axis][i] * hm[j];
      const double mji = X[axis][j] * hm[i];
      s += 0.5 * (mji + mij) * sig[i][j];
    }
  return s;
}

/* energy and gradient [8][3] of one state; model 0 = eight corners, 1 = averaged matrix */
static double state_energy(int model, const double verts[8][3], const double* h0, const double* C,
                           double g[8][3]) {
  double hi[9];
  const double det = inv3(h0, hi);
  double M[8][3][3];
  for (int a = 0; a < 8; ++a)
    for (int d = 0; d < 3; ++d)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int v = 0; v < 8; ++v) s += g_mult[a][d][v] * verts[v][j];
        M[a][d][j] = s;
      }
  memset(g, 0, sizeof(double) * 24);
  if (model == 0) {
    double e = 0.0;
    for (int a = 0; a < 8; ++a) {
      double X[3][3], eps[3][3], sig[3][3];
      strain_of(M[a], hi, X, eps);
      e += stress_of(C, eps, sig);
      for (int v = 0; v < 8; ++v) {
        const double mu[3] = {g_mult[a][0][v], g_mult[a][1][v], g_mult[a][2][v]};
        if (mu[0] == 0.0 && mu[1] == 0.0 && mu[2] == 0.0) continue;
        for (int axis = 0; axis < 3; ++axis)
          g[v][axis] += 0.125 * dstrain_dot(X, hi, mu, axis, sig);
      }
    }
    for (int v = 0; v < 8; ++v)
      for (int axis = 0; axis < 3; ++axis) g[v][axis] *= det;
    return 0.125 * e * det;
  }
  double H[3][3];
  for (int d = 0; d < 3; ++d)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int a = 0; a < 8; ++a) s += M[a][d][j];
      H[d][j] = 0.125 * s;
    }
  double X[3][3], eps[3][3], sig[3][3];
  strain_of(H, hi, X, eps);
  const double u = stress_of(C, eps, sig);
  for (int v = 0; v < 8; ++v) {
    const double mu[3] = {0.25 * g_mult[v][0][v], 0.25 * g_mult[v][1][v], 0.25 * g_mult[v][2][v]};
    for (int axis = 0; axis < 3; ++axis) g[v][axis] = det * dstrain_dot(X, hi, mu, axis, sig);
  }
  return u * det;
}

/* Energy, gradient and virial of a whole system.
 *   pos [N][3], rvecs [nvec][3] (nvec = 0 or 3), surrounding_nodes [C][8],
 *   surrounding_cells [N][8] (-1 = none), wrap [C] (bit d: last layer along d),
 *   cell_row [C] -> parameter row r < nrows with nstates[r] <= smax states,
 *   h0 [nrows][smax][9] (rows = cell vectors), c0 [nrows][smax][81], free [nrows][smax],
 *   temp [nrows] (effective temperature in K), kB in hartree per K.
 * gpos [N][3] and vtens [9] may be NULL. Returns 0, or -1 when out of memory. */
int mmo_compute(int model, int64_t C, int64_t N, const double* pos, const double* rvecs,
                int nvec, const int64_t* surrounding_nodes, const int64_t* surrounding_cells,
                const uint8_t* wrap, const int32_t* cell_row, int nrows, int smax,
                const int32_t* nstates, const double* h0, const double* c0, const double* free_e,
                const double* temp, double kB, double* energy, double* gpos, double* vtens) {
  static int ready = 0;
  if (!ready) {
    init_mult();
    ready = 1;
  }
  (void)nrows;
  double rv[3][3] = {{0}};
  for (int d = 0; d < nvec && d < 3; ++d)
    for (int j = 0; j < 3; ++j) rv[d][j] = rvecs[d * 3 + j];
  double* ecell = (double*)malloc(sizeof(double) * (C > 0 ? C : 1));
  double* gcell = (double*)malloc(sizeof(double) * 24 * (C > 0 ? C : 1));
  double* vcell = (double*)malloc(sizeof(double) * 9 * (C > 0 ? C : 1));
  if (!ecell || !gcell || !vcell) {
    free(ecell);
    free(gcell);
    free(vcell);
    return -1;
  }
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < C; ++c) {
    const int64_t* sn = surrounding_nodes + c * 8;
    double verts[8][3];
    const double* r0 = pos + sn[0] * 3;
    for (int k = 0; k < 8; ++k)
      for (int j = 0; j < 3; ++j) {
        double d = pos[sn[k] * 3 + j] - r0[j];
        for (int ax = 0; ax < 3; ++ax)
          if (nvec > 0 && ((wrap[c] >> ax) & 1) && kOff[k][ax]) d += rv[ax][j];
        verts[k][j] = k == 0 ? r0[j] : r0[j] + d;
      }
    const int r = cell_row[c];
    const int S = nstates[r];
    const double kT = kB * temp[r];
    double es[8], gs[8][8][3];
    double emin = INFINITY;
    for (int s = 0; s < S && s < 8; ++s) {
      es[s] = state_energy(model, verts, h0 + ((int64_t)r * smax + s) * 9,
                           c0 + ((int64_t)r * smax + s) * 81, gs[s]) +
              free_e[(int64_t)r * smax + s];
      if (es[s] < emin) emin = es[s];
    }
    double w[8], wsum = 0.0;
    for (int s = 0; s < S; ++s) {
      w[s] = exp(-(es[s] - emin) / kT);
      wsum += w[s];
    }
    ecell[c] = emin - kT * log(wsum);
    double* g = gcell + c * 24;
    for (int i = 0; i < 24; ++i) g[i] = 0.0;
    for (int s = 0; s < S; ++s)
      for (int v = 0; v < 8; ++v)
        for (int j = 0; j < 3; ++j) g[v * 3 + j] += (w[s] / wsum) * gs[s][v][j];
    double* vt = vcell + c * 9;
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 3; ++l) {
        double t = 0.0;
        for (int v = 0; v < 8; ++v) t += g[v * 3 + k] * verts[v][l];
        vt[k * 3 + l] = t;
      }
  }
  /* sums in a fixed order: blocks of 4096 cells, then the blocks in order */
  const int64_t B = 4096;
  const int64_t nb = (C + B - 1) / B;
  double* part = (double*)calloc((size_t)(nb > 0 ? nb : 1) * 10, sizeof(double));
  if (!part) {
    free(ecell);
    free(gcell);
    free(vcell);
    return -1;
  }
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < nb; ++b) {
    double acc[10] = {0};
    for (int64_t c = b * B; c < C && c < (b + 1) * B; ++c) {
      acc[0] += ecell[c];
      for (int i = 0; i < 9; ++i) acc[1 + i] += vcell[c * 9 + i];
    }
    for (int i = 0; i < 10; ++i) part[b * 10 + i] = acc[i];
  }
  double tot[10] = {0};
  for (int64_t b = 0; b < nb; ++b)
    for (int i = 0; i < 10; ++i) tot[i] += part[b * 10 + i];
  *energy = tot[0];
  if (vtens)
    for (int i = 0; i < 9; ++i) vtens[i] = tot[1 + i];
  if (gpos) {
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < N; ++n) {
      double acc[3] = {0.0, 0.0, 0.0};
      for (int k = 0; k < 8; ++k) {
        const int64_t c = surrounding_cells[n * 8 + k];
        if (c < 0) continue;
        for (int j = 0; j < 3; ++j) acc[j] += gcell[c * 24 + k * 3 + j];
      }
      for (int j = 0; j < 3; ++j) gpos[n * 3 + j] = acc[j];
    }
  }
  free(part);
  free(ecell);
  free(gcell);
  free(vcell);
  return 0;
}

int mmo_num_threads(void) {
#ifdef _OPENMP
  extern int omp_get_max_threads(void);
  return omp_get_max_threads();
#else
  return 1;
#endif
}
