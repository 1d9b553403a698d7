/*
 * psb_oracle.c -- CPU restatement (FP64) of PSI-BOIL's pressure-Poisson path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never shipped
 * behind it.  Compile with -O2 -ffp-contract=off (no fused multiply-add) so
 * that every per-cell expression rounds like the reference built with g++.
 *
 * How each function is pinned (tests/test_oracle*.py):
 *   ora_spmv, ora_resid          bit-exact against oracle/_ref (the
 *                                reference's own Scalar operators) on the
 *                                golden fixtures; dense-matrix brute force;
 *                                A*1 = 0 on pure-Neumann rows; symmetry.
 *   ora_dot, ora_dot_voldiv_avg  closed form (math.fsum) on random vectors.
 *   ora_cg, ora_bicgs            residual histories and iteration counts of
 *                                the reference's CG::solve / BiCGS::solve
 *                                (golden fixtures); dense pseudo-inverse
 *                                solution on tiny grids.
 *   ora_cg_prec (IC0, IC2)       x after 1, 5 iterations and at 1e-8, counts
 *                                and residuals of the reference's CG with
 *                                Prec::ic0 / ic2 (prec fixtures), bit for bit.
 *   ora_set_kcycle (K-cycle)     no counterpart in the reference: an
 *                                independent numpy restatement on the
 *                                oracle's own levels, homogeneity, and the
 *                                orthogonality of the two steps' residual to
 *                                both search directions.
 *   ora_cg_prec (RBSGS),
 *   ora_prec_apply               no counterpart in the reference: z = M^-1 r
 *                                against M = (D - L) D^-1 (D - U) assembled
 *                                densely in red-black order (the other order
 *                                fails it), M = A + L D^-1 U, symmetry.
 *   ora_jacobi                   reference Jacobi::solve history (fixtures);
 *                                closed-form iteration on a diagonal matrix.
 *   ora_levels                   reference level sizes (fixtures) and the
 *                                factor() known answers.
 *   ora_coarsen_flag/_system,
 *   ora_restriction,
 *   ora_interpolation            reference level matrices / restricted RHS
 *                                (fixtures) and dense P^T A P, P^T(f-A phi)
 *                                with explicit P on 8^3 -> 4^3.
 *   ora_ac_vcycle, ora_ac_cycle  reference AC::vcycle residual history and
 *                                phi (fixtures); the F, W and full-cycle
 *                                shapes, linf control, divergence and
 *                                min-cycle controls: residual histories
 *                                and phi of the reference's own AC::cycle
 *                                (cycles fixtures), bit for bit.
 *   ora_discretize, ora_rhs      reference Pressure::discretize / update_rhs
 *                                on the same grid metrics and density
 *                                (fixtures); closed form on uniform grids.
 *   ora_project                  the identity rhs(project(u, p)) = rhs(u) -
 *                                dt A p on every fluid cell for random p,
 *                                with ora_rhs / ora_discretize / ora_spmv on
 *                                grids stretched in x, y and z; divergence
 *                                after a converged solve = its residual.
 *   ora_ac_coarse_dense,
 *   ora_ac_precond, ora_acpcg,
 *   ora_pcg_rel                  not in the reference: dense pseudo-inverse
 *                                solve on tiny grids, symmetry of M^-1
 *                                (u.M^-1 v = v.M^-1 u), A-orthogonality.
 *   ora_set_flexible             not in the reference: with the flexible
 *                                beta every new search direction is
 *                                A-orthogonal to the last one
 *                                (ora_set_orth_record) under the K-cycle's
 *                                varying preconditioner, which the
 *                                Fletcher-Reeves beta fails; with a fixed
 *                                preconditioner both agree.
 *   ora_zline_solve,
 *   ora_ac_precond_s             not in the reference: per-column dense
 *                                tridiagonal solves; symmetry of M^-1 with
 *                                the z-line smoother.
 */
#include "psb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j, k, ny, nz) (((size_t)(i) * (size_t)(ny) + (size_t)(j)) * (size_t)(nz) + (size_t)(k))

static const double PICO = 1.0e-12; /* Src/Global/global_constants.cpp:17 */
static const double ATTO = 1.0e-18; /* Src/Global/global_constants.cpp:19 */
static const double UNREAL = 1.0e+24; /* boil::unreal = yotta */

/* ------------------------------------------------------------------------ */
/* generator                                                                 */
/* ------------------------------------------------------------------------ */
uint64_t ora_splitmix64(uint64_t *state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double ora_u01(uint64_t *state) {
  return (double)(ora_splitmix64(state) >> 11) * (1.0 / 9007199254740992.0);
}

void ora_gen_spheres(uint64_t seed, int nsph, double rmin, double rmax,
                     double Lx, double Ly, double Lz, double *out) {
  uint64_t st = seed;
  for (int s = 0; s < nsph; s++) {
    double r = rmin + (rmax - rmin) * ora_u01(&st);
    double cx = r + (Lx - 2.0 * r) * ora_u01(&st);
    double cy = r + (Ly - 2.0 * r) * ora_u01(&st);
    double cz = r + (Lz - 2.0 * r) * ora_u01(&st);
    out[4 * s + 0] = cx;
    out[4 * s + 1] = cy;
    out[4 * s + 2] = cz;
    out[4 * s + 3] = r;
  }
}

/* Src/Grid/grid1d_correct_boundaries.cpp: BndGrid::wall() keeps the ghost
 * nodes on the wall, so the ghost cell centre is the wall node itself. */
void ora_metrics(int n, const double *xn, double *xc, double *dxc, double *dxw,
                 double *dxe) {
  for (int i = 0; i < n; i++) {
    xc[i] = 0.5 * (xn[i + 1] + xn[i]);
    dxc[i] = xn[i + 1] - xn[i];
  }
  for (int i = 0; i < n; i++) {
    double xm = (i == 0) ? 0.5 * (xn[0] + xn[0]) : xc[i - 1];
    double xp = (i == n - 1) ? 0.5 * (xn[n] + xn[n]) : xc[i + 1];
    dxw[i] = xc[i] - xm;
    dxe[i] = xp - xc[i];
  }
}

void ora_density(int nx, int ny, int nz, const double *xc, const double *yc,
                 const double *zc, int nsph, const double *sph, double rho_in,
                 double rho_out, double *rho) {
  const int gy = ny + 2, gz = nz + 2;
  for (int i = -1; i <= nx; i++)
    for (int j = -1; j <= ny; j++)
      for (int k = -1; k <= nz; k++) {
        int ii = i < 0 ? 0 : (i >= nx ? nx - 1 : i);
        int jj = j < 0 ? 0 : (j >= ny ? ny - 1 : j);
        int kk = k < 0 ? 0 : (k >= nz ? nz - 1 : k);
        int inside = 0;
        for (int s = 0; s < nsph && !inside; s++) {
          double dx = xc[ii] - sph[4 * s + 0];
          double dy = yc[jj] - sph[4 * s + 1];
          double dz = zc[kk] - sph[4 * s + 2];
          if (sqrt(dx * dx + dy * dy + dz * dz) < sph[4 * s + 3]) inside = 1;
        }
        rho[((size_t)(i + 1) * gy + (j + 1)) * gz + (k + 1)] =
            inside ? rho_in : rho_out;
      }
}

void ora_discretize(int nx, int ny, int nz, const double *dxc, const double *dxw,
                    const double *dxe, const double *dyc, const double *dys,
                    const double *dyn, const double *dzc, const double *dzb,
                    const double *dzt, const double *rho, const uint8_t *off,
                    ora_mat *A, double *ci, double *dV) {
  ora_discretize_bnd(nx, ny, nz, dxc, dxw, dxe, dyc, dys, dyn, dzc, dzb, dzt, rho, off, NULL,
                     NULL, A, ci, dV, NULL);
}

void ora_discretize_bnd(int nx, int ny, int nz, const double *dxc, const double *dxw,
                        const double *dxe, const double *dyc, const double *dys,
                        const double *dyn, const double *dzc, const double *dzb,
                        const double *dzt, const double *rho, const uint8_t *off,
                        const int *btype, const double *bval, ora_mat *A, double *ci,
                        double *dV, double *fbnd) {
  const int gy = ny + 2, gz = nz + 2;
#define RHO(i, j, k) rho[((size_t)((i) + 1) * gy + ((j) + 1)) * gz + ((k) + 1)]
  const size_t N = (size_t)nx * ny * nz;
  double *csum = off ? (double *)malloc(N * sizeof(double)) : NULL;
  A->nx = nx; A->ny = ny; A->nz = nz;
  /* neighbouring coefficients, pressure_discretize.cpp:24-57 */
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        size_t P = IDX(i, j, k, ny, nz);
        double rp = RHO(i, j, k);
        double a_x = dyc[j] * dzc[k];
        double a_y = dxc[i] * dzc[k];
        double a_z = dxc[i] * dyc[j];
        A->w[P] = a_x / dxw[i] / (0.5 * (rp + RHO(i - 1, j, k)));
        A->e[P] = a_x / dxe[i] / (0.5 * (RHO(i + 1, j, k) + rp));
        A->s[P] = a_y / dys[j] / (0.5 * (rp + RHO(i, j - 1, k)));
        A->n[P] = a_y / dyn[j] / (0.5 * (RHO(i, j + 1, k) + rp));
        A->b[P] = a_z / dzb[k] / (0.5 * (rp + RHO(i, j, k - 1)));
        A->t[P] = a_z / dzt[k] / (0.5 * (RHO(i, j, k + 1) + rp));
        /* central coefficient, :65-67 */
        A->c[P] = A->w[P] + A->e[P] + A->s[P] + A->n[P] + A->b[P] + A->t[P];
        if (csum) csum[P] = A->c[P];
      }
  if (off) {
    /* voxel body: zero the couplings of off cells on both sides (:93-106) */
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          size_t P = IDX(i, j, k, ny, nz);
          if (!off[P]) continue;
          A->w[P] = A->e[P] = A->s[P] = A->n[P] = A->b[P] = A->t[P] = 0.0;
          if (i > 0) A->e[IDX(i - 1, j, k, ny, nz)] = 0.0;
          if (i < nx - 1) A->w[IDX(i + 1, j, k, ny, nz)] = 0.0;
          if (j > 0) A->n[IDX(i, j - 1, k, ny, nz)] = 0.0;
          if (j < ny - 1) A->s[IDX(i, j + 1, k, ny, nz)] = 0.0;
          if (k > 0) A->t[IDX(i, j, k - 1, ny, nz)] = 0.0;
          if (k < nz - 1) A->b[IDX(i, j, k + 1, ny, nz)] = 0.0;
        }
    /* central coefficient of cells that are not off (:112-117) */
    for (size_t P = 0; P < N; P++)
      A->c[P] = off[P] ? csum[P]
                       : A->w[P] + A->e[P] + A->s[P] + A->n[P] + A->b[P] + A->t[P];
  }
  /* boundary sides (centered_system_bnd.cpp:5-61, 119-160): first the
   * Neumann group (wall, symmetry, outlet, neumann; btype 0) in the order
   * imin, imax, jmin, jmax, kmin, kmax -- the side's coupling moves out of c,
   * value * dS joins fbnd -- then the Dirichlet group (dirichlet, inlet;
   * btype 1), whose coupling stays in c and joins fbnd times the value the
   * ghost holds.  btype NULL: walls everywhere. */
  if (fbnd)
    for (size_t P = 0; P < N; P++) fbnd[P] = 0.0;
  for (int grp = 0; grp < 2; grp++)
    for (int side = 0; side < 6; side++) {
      const int t = btype ? btype[side] : 0;
      const double v = bval ? bval[side] : 0.0;
      if (t != grp) continue;
      const int ax = side / 2, hi = side & 1;
      const int n1 = ax == 0 ? ny : nx, n2 = ax == 2 ? ny : nz;
      for (int p = 0; p < n1; p++)
        for (int q = 0; q < n2; q++) {
          int i = 0, j = 0, k = 0;
          if (ax == 0) { i = hi ? nx - 1 : 0; j = p; k = q; }
          if (ax == 1) { i = p; j = hi ? ny - 1 : 0; k = q; }
          if (ax == 2) { i = p; j = q; k = hi ? nz - 1 : 0; }
          const size_t P = IDX(i, j, k, ny, nz);
          double *cp = ax == 0 ? (hi ? A->e : A->w) : ax == 1 ? (hi ? A->n : A->s)
                                                              : (hi ? A->t : A->b);
          const double dS = ax == 0 ? dyc[j] * dzc[k] : ax == 1 ? dxc[i] * dzc[k]
                                                                : dxc[i] * dyc[j];
          if (grp == 0) {
            if (fbnd) fbnd[P] += v * dS;
            A->c[P] -= cp[P];
          } else if (fbnd) {
            fbnd[P] += v * cp[P];
          }
          cp[P] = 0.0;
        }
    }
  /* ci = 1/c before the off-cell scaling (:127), then c *= pico (:132-148) */
  for (size_t P = 0; P < N; P++) {
    if (ci) ci[P] = 1.0 / A->c[P];
    if (off && off[P]) A->c[P] *= PICO;
  }
  if (dV)
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++)
          dV[IDX(i, j, k, ny, nz)] = dxc[i] * dyc[j] * dzc[k];
  free(csum);
#undef RHO
}

void ora_face_velocity(uint64_t seed, int nx, int ny, int nz, double *u,
                       double *v, double *w) {
  uint64_t st = seed;
  for (int i = 0; i <= nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        double val = 2.0 * ora_u01(&st) - 1.0;
        u[((size_t)i * ny + j) * nz + k] = (i == 0 || i == nx) ? 0.0 : val;
      }
  for (int i = 0; i < nx; i++)
    for (int j = 0; j <= ny; j++)
      for (int k = 0; k < nz; k++) {
        double val = 2.0 * ora_u01(&st) - 1.0;
        v[((size_t)i * (ny + 1) + j) * nz + k] = (j == 0 || j == ny) ? 0.0 : val;
      }
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k <= nz; k++) {
        double val = 2.0 * ora_u01(&st) - 1.0;
        w[((size_t)i * ny + j) * (nz + 1) + k] = (k == 0 || k == nz) ? 0.0 : val;
      }
}

/* pressure_update_rhs.cpp:27-52 with dti = 1, off cells zero (:56-57) */
void ora_rhs(int nx, int ny, int nz, const double *dxc, const double *dyc,
             const double *dzc, const double *u, const double *v, const double *w,
             const uint8_t *off, int zero_mean, double *b) {
  ora_rhs_bnd(nx, ny, nz, dxc, dyc, dzc, u, v, w, off, NULL, zero_mean, b);
}

void ora_rhs_bnd(int nx, int ny, int nz, const double *dxc, const double *dyc,
                 const double *dzc, const double *u, const double *v, const double *w,
                 const uint8_t *off, const double *fbnd, int zero_mean, double *b) {
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        size_t P = IDX(i, j, k, ny, nz);
        double a_x = dyc[j] * dzc[k] * 1.0;
        double a_y = dxc[i] * dzc[k] * 1.0;
        double a_z = dxc[i] * dyc[j] * 1.0;
        double f = 0.0;
        f += a_x * u[((size_t)i * ny + j) * nz + k];
        f -= a_x * u[((size_t)(i + 1) * ny + j) * nz + k];
        f += a_y * v[((size_t)i * (ny + 1) + j) * nz + k];
        f -= a_y * v[((size_t)i * (ny + 1) + j + 1) * nz + k];
        f += a_z * w[((size_t)i * ny + j) * (nz + 1) + k];
        f -= a_z * w[((size_t)i * ny + j) * (nz + 1) + k + 1];
        if (off && off[P]) f = 0.0;
        if (fbnd) f += 0.0 + fbnd[P]; /* fnew += fext + fbnd (:135), fext 0 */
        b[P] = f;
      }
  if (zero_mean) {
    const size_t N = (size_t)nx * ny * nz;
    double sum = 0.0;
    long cnt = 0;
    for (size_t P = 0; P < N; P++)
      if (!(off && off[P])) { sum += b[P]; cnt++; }
    if (cnt > 0) {
      double mean = sum / (double)cnt;
      for (size_t P = 0; P < N; P++)
        if (!(off && off[P])) b[P] -= mean;
    }
  }
}

/* Momentum::project (Src/Equation/Staggered/Momentum/momentum_project.cpp:
 * 10-76): the interior faces (Staggered::set_indices drops the wall faces,
 * staggered_set_indices.cpp:9-19) take the pressure-gradient correction with
 * the face density 0.5 (rho_P + rho_nb) (property_value.cpp:18-32) and the
 * centre distance dxw / dys / dzb (Vector::dxc_staggered, vector.h:388).
 * Without a body (off == NULL) u += (p_W - p_P) / rho * dt / dxw (:29-33);
 * with one, only faces between two fluid cells, as (p_W - p_P) * dt / rho /
 * dxw (:50-55).  mask: bit 0 u, bit 1 v, bit 2 w (ifull / jfull / kfull).   */
void ora_project(int nx, int ny, int nz, const double *dxw, const double *dys,
                 const double *dzb, const double *rho, const uint8_t *off,
                 const double *p, double dt, int mask, double *u, double *v,
                 double *w) {
  const int gy = ny + 2, gz = nz + 2;
#define RHO(i, j, k) rho[((size_t)((i) + 1) * gy + ((j) + 1)) * gz + ((k) + 1)]
  if (mask & 1)
    for (int i = 1; i < nx; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          size_t P = IDX(i, j, k, ny, nz), W = IDX(i - 1, j, k, ny, nz);
          double r = 0.5 * (RHO(i, j, k) + RHO(i - 1, j, k));
          double *f = &u[((size_t)i * ny + j) * nz + k];
          if (!off) *f += (p[W] - p[P]) / r * dt / dxw[i];
          else if (!off[P] && !off[W]) *f += (p[W] - p[P]) * dt / r / dxw[i];
        }
  if (mask & 2)
    for (int i = 0; i < nx; i++)
      for (int j = 1; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          size_t P = IDX(i, j, k, ny, nz), S = IDX(i, j - 1, k, ny, nz);
          double r = 0.5 * (RHO(i, j, k) + RHO(i, j - 1, k));
          double *f = &v[((size_t)i * (ny + 1) + j) * nz + k];
          if (!off) *f += (p[S] - p[P]) / r * dt / dys[j];
          else if (!off[P] && !off[S]) *f += (p[S] - p[P]) * dt / r / dys[j];
        }
  if (mask & 4)
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 1; k < nz; k++) {
          size_t P = IDX(i, j, k, ny, nz), B = IDX(i, j, k - 1, ny, nz);
          double r = 0.5 * (RHO(i, j, k) + RHO(i, j, k - 1));
          double *f = &w[((size_t)i * ny + j) * (nz + 1) + k];
          if (!off) *f += (p[B] - p[P]) / r * dt / dzb[k];
          else if (!off[P] && !off[B]) *f += (p[B] - p[P]) * dt / r / dzb[k];
        }
#undef RHO
}

/* kappa_avei / avej / avek (pressure_ghost.cpp:192-213): the harmonic mean
 * of two curvatures of one sign, else their arithmetic mean */
static double kappa_face(double a, double b) {
  double k = 0.5 * (a + b);
  if (a * b > 0.0) k = 2.0 * a * b / (a + b);
  return k;
}

/* Pressure::ghost (pressure_ghost.cpp:10-190): for each face of a cell that
 * the colour function crosses at c = 0.5 (one side below, the other above),
 * fext += iphase dS sigma kappa_face / (rho_face d), iphase = -1 when the
 * cell's c is below 0.5.  Faces in the order w, e, s, n, b, t; rho_face =
 * 0.5 (rho_nb + rho_P) as PropertyMix::value(Comp, ...) forms it
 * (property_value.cpp:59-83); d = dxw, dxe, dys, dys (the north face uses
 * dys, :74), dzb, dzt.  c, kappa, rho with their ghost layer (as
 * ora_discretize's rho); solid cells (off) are skipped (:96).               */
void ora_ghost(int nx, int ny, int nz, const double *dxc, const double *dxw,
               const double *dxe, const double *dyc, const double *dys,
               const double *dzc, const double *dzb, const double *dzt,
               const double *c, const double *kappa, const double *rho, double sigma,
               const uint8_t *off, double *fext) {
  const int gy = ny + 2, gz = nz + 2;
#define G3(a, i, j, k) a[((size_t)((i) + 1) * gy + ((j) + 1)) * gz + ((k) + 1)]
  const double cint = 0.5;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        const size_t P = IDX(i, j, k, ny, nz);
        if (off && off[P]) continue;
        const double cc = G3(c, i, j, k);
        const double iphase = cc < cint ? -1.0 : 1.0;
        double f = fext[P];
        /* one face: neighbour (a, b, cn) offsets, face area, distance; the
         * face's density and curvature pair are (nb, P) for the low side and
         * (P, nb) for the high side, as the reference indexes them */
#define FACE(di, dj, dk, area, dd)                                                      \
  do {                                                                                  \
    const double cn = G3(c, i + (di), j + (dj), k + (dk));                              \
    if ((cc - cint) * (cn - cint) < 0.0) {                                              \
      const int lo = (di) + (dj) + (dk) < 0;                                            \
      const double r0 = lo ? G3(rho, i + (di), j + (dj), k + (dk)) : G3(rho, i, j, k);  \
      const double r1 = lo ? G3(rho, i, j, k) : G3(rho, i + (di), j + (dj), k + (dk));  \
      const double k0 = lo ? G3(kappa, i + (di), j + (dj), k + (dk)) : G3(kappa, i, j, k); \
      const double k1 = lo ? G3(kappa, i, j, k) : G3(kappa, i + (di), j + (dj), k + (dk)); \
      const double rf = 0.5 * (r1 + r0);                                                \
      f += iphase * (area) * sigma * kappa_face(k0, k1) / (rf * (dd));                  \
    }                                                                                   \
  } while (0)
        FACE(-1, 0, 0, dyc[j] * dzc[k], dxw[i]);
        FACE(1, 0, 0, dyc[j] * dzc[k], dxe[i]);
        FACE(0, -1, 0, dxc[i] * dzc[k], dys[j]);
        FACE(0, 1, 0, dxc[i] * dzc[k], dys[j]);
        FACE(0, 0, -1, dxc[i] * dyc[j], dzb[k]);
        FACE(0, 0, 1, dxc[i] * dyc[j], dzt[k]);
#undef FACE
        fext[P] = f;
      }
#undef G3
}

/* ------------------------------------------------------------------------ */
/* vector kernels                                                            */
/* ------------------------------------------------------------------------ */
static inline double nbx(const double *x, int i, int j, int k, int nx, int ny,
                         int nz) {
  if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) return 0.0;
  return x[IDX(i, j, k, ny, nz)];
}

/* Src/Field/Scalar/scalar_operators.cpp:49-72 (four passes per cell, same
 * grouping of the products within each pass) */
void ora_spmv(const ora_mat *A, const double *x, double *y) {
  const int nx = A->nx, ny = A->ny, nz = A->nz;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        size_t P = IDX(i, j, k, ny, nz);
        double v = A->c[P] * x[P];
        v -= (A->w[P] * nbx(x, i - 1, j, k, nx, ny, nz) +
              A->e[P] * nbx(x, i + 1, j, k, nx, ny, nz));
        v -= (A->s[P] * nbx(x, i, j - 1, k, nx, ny, nz) +
              A->n[P] * nbx(x, i, j + 1, k, nx, ny, nz));
        v -= (A->b[P] * nbx(x, i, j, k - 1, nx, ny, nz) +
              A->t[P] * nbx(x, i, j, k + 1, nx, ny, nz));
        y[P] = v;
      }
}

/* Src/Field/Scalar/scalar_operators.cpp:22-46 */
void ora_resid(const ora_mat *A, const double *b, const double *x, double *r) {
  const int nx = A->nx, ny = A->ny, nz = A->nz;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        size_t P = IDX(i, j, k, ny, nz);
        double v = b[P] - A->c[P] * x[P];
        v += A->w[P] * nbx(x, i - 1, j, k, nx, ny, nz) +
             A->e[P] * nbx(x, i + 1, j, k, nx, ny, nz);
        v += A->s[P] * nbx(x, i, j - 1, k, nx, ny, nz) +
             A->n[P] * nbx(x, i, j + 1, k, nx, ny, nz);
        v += A->b[P] * nbx(x, i, j, k - 1, nx, ny, nz) +
             A->t[P] * nbx(x, i, j, k + 1, nx, ny, nz);
        r[P] = v;
      }
}

/* Src/Field/Scalar/scalar.h:296-298 */
/* Summation order of the dot products: 0 (default) sequential as the
 * reference; b > 0 sums blocks of b terms and then the block sums in order,
 * as a blocked parallel reduction does.  Test infrastructure: measures how
 * much an algorithm's result depends on the order of its sums. */
static size_t sum_block = 0;
void ora_set_sum_block(int b) { sum_block = b > 0 ? (size_t)b : 0; }

double ora_dot(size_t n, const double *x, const double *y) {
  if (sum_block) {
    double d = 0.0;
    for (size_t q = 0; q < n; q += sum_block) {
      double t = 0.0;
      for (size_t P = q; P < n && P < q + sum_block; P++) t += x[P] * y[P];
      d += t;
    }
    return d;
  }
  double d = 0.0;
  for (size_t P = 0; P < n; P++) d += x[P] * y[P];
  return d;
}

/* Src/Field/Scalar/scalar.h:317-326 */
double ora_dot_voldiv_avg(size_t n, const double *x, const double *y,
                          const double *dV) {
  double d = 0.0;
  if (sum_block) {
    for (size_t q = 0; q < n; q += sum_block) {
      double t = 0.0;
      for (size_t P = q; P < n && P < q + sum_block; P++) t += x[P] * y[P] / dV[P] / dV[P];
      d += t;
    }
  } else {
    for (size_t P = 0; P < n; P++) d += x[P] * y[P] / dV[P] / dV[P];
  }
  return d / (double)n / (double)n;
}

static size_t mat_n(const ora_mat *A) { return (size_t)A->nx * A->ny * A->nz; }

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* staleness window shared by CG and BiCGS (cg.cpp:36-41,138-160) */
static int stale_check(double *win, int cnt, double res) {
  for (int q = 0; q < cnt; q++)
    if (res < win[q]) {
      memmove(win, win + 1, (size_t)(cnt - 1) * sizeof(double));
      win[cnt - 1] = res;
      return 0;
    }
  return 1;
}

/* ------------------------------------------------------------------------ */
/* preconditioners of CG (Src/Solver/Preconditioner, selected by            */
/* linear.cpp:31-34): Diagonal, IncompleteCholesky0 / 2, and the red-black  */
/* symmetric Gauss-Seidel that the library runs in place of the sequential  */
/* incomplete-Cholesky sweeps                                               */
/* ------------------------------------------------------------------------ */
typedef struct {
  int kind;
  const ora_mat *A;
  double *Mc, *Mw, *Ms, *Mb, *Mtw, *Mts;
} ora_prec;

/* IncompleteCholesky0 / 2 :: form (preconditioner_form.cpp:16-80) on the
 * compact interior: the factor's couplings the reference writes into ghost
 * cells multiply ghost values of z and are left out */
static void prec_form(ora_prec *M, const ora_mat *A, int kind) {
  const int nx = A->nx, ny = A->ny, nz = A->nz;
  const size_t N = mat_n(A);
  memset(M, 0, sizeof(*M));
  M->kind = kind;
  M->A = A;
  M->Mc = dalloc(N);
  if (kind == ORA_PREC_DI || kind == ORA_PREC_RBSGS) {
    for (size_t P = 0; P < N; P++) M->Mc[P] = 1.0 / A->c[P];
    return;
  }
  const int ic2 = kind == ORA_PREC_IC2;
  M->Mw = dalloc(N); M->Ms = dalloc(N); M->Mb = dalloc(N);
  M->Mtw = dalloc(N); M->Mts = dalloc(N);
  double *Mc = M->Mc, *Mw = M->Mw, *Ms = M->Ms, *Mb = M->Mb, *Mtw = M->Mtw, *Mts = M->Mts;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        const size_t P = IDX(i, j, k, ny, nz);
        double sum = A->c[P];
        sum = sum - Mw[P] * Mw[P];
        if (ic2) sum = sum - Mtw[P] * Mtw[P];
        sum = sum - Ms[P] * Ms[P];
        if (ic2) sum = sum - Mts[P] * Mts[P];
        sum = sum - Mb[P] * Mb[P];
        sum = sqrt(sum);
        Mc[P] = sum;
        if (i + 1 < nx) {
          const size_t E = IDX(i + 1, j, k, ny, nz);
          Mw[E] = -A->w[E] / Mc[P];
          if (ic2 && k > 0) {
            const size_t EB = IDX(i + 1, j, k - 1, ny, nz);
            Mtw[EB] = -Mw[EB] * Mb[P] / Mc[P];
          }
        }
        if (j + 1 < ny) {
          const size_t Nn = IDX(i, j + 1, k, ny, nz);
          Ms[Nn] = -A->s[Nn] / Mc[P];
          if (ic2 && k > 0) {
            const size_t NB = IDX(i, j + 1, k - 1, ny, nz);
            Mts[NB] = -Ms[NB] * Mb[P] / Mc[P];
          }
        }
        if (k + 1 < nz) {
          const size_t T = P + 1;
          if (ic2)
            Mb[T] = (-A->b[T] - Mw[T] * Mtw[P] - Ms[T] * Mts[P]) / Mc[P];
          else
            Mb[T] = -A->b[T] / Mc[P];
        }
      }
}

static void prec_free(ora_prec *M) {
  free(M->Mc); free(M->Mw); free(M->Ms); free(M->Mb); free(M->Mtw); free(M->Mts);
}

/* one colour of the red-black symmetric Gauss-Seidel sweep on the cells with
 * (i + j + k) % 2 == colour, in place:  z_P = z_P + (sum a_nb z_nb) / c_P,
 * the neighbour sum grouped as (W + E) + (S + N) + (B + T) */
static void rbsgs_colour(const ora_mat *A, double *z, int colour) {
  const int nx = A->nx, ny = A->ny, nz = A->nz;
  const size_t sI = (size_t)ny * nz, sJ = (size_t)nz;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        if (((i + j + k) & 1) != colour) continue;
        const size_t P = IDX(i, j, k, ny, nz);
        const double xw = i > 0 ? z[P - sI] : 0.0, xe = i < nx - 1 ? z[P + sI] : 0.0;
        const double xs = j > 0 ? z[P - sJ] : 0.0, xn = j < ny - 1 ? z[P + sJ] : 0.0;
        const double xb = k > 0 ? z[P - 1] : 0.0, xt = k < nz - 1 ? z[P + 1] : 0.0;
        const double s1 = A->w[P] * xw + A->e[P] * xe;
        const double s2 = A->s[P] * xs + A->n[P] * xn;
        const double s3 = A->b[P] * xb + A->t[P] * xt;
        z[P] = z[P] + ((s1 + s2) + s3) / A->c[P];
      }
}

/* z = M^-1 r */
static void prec_solve(const ora_prec *M, double *z, const double *r) {
  const ora_mat *A = M->A;
  const int nx = A->nx, ny = A->ny, nz = A->nz;
  const size_t N = mat_n(A);
  if (M->kind == ORA_PREC_DI) {  /* preconditioner_solve.cpp:4-11 */
    for (size_t P = 0; P < N; P++) z[P] = r[P] * M->Mc[P];
    return;
  }
  if (M->kind == ORA_PREC_RBSGS) {
    /* M = (D - L) D^-1 (D - U) in red-black order (even cells first):
     * z = D^-1 r, odd cells from the even ones, then even cells from the
     * odd ones */
    for (size_t P = 0; P < N; P++) z[P] = r[P] / A->c[P];
    rbsgs_colour(A, z, 1);
    rbsgs_colour(A, z, 0);
    return;
  }
  /* IncompleteCholesky0 / 2 :: solve (preconditioner_solve.cpp:14-73) */
  const int ic2 = M->kind == ORA_PREC_IC2;
  const double *Mc = M->Mc, *Mw = M->Mw, *Ms = M->Ms, *Mb = M->Mb, *Mtw = M->Mtw,
               *Mts = M->Mts;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        const size_t P = IDX(i, j, k, ny, nz);
        double sum = r[P];
        if (i > 0) sum = sum - Mw[P] * z[IDX(i - 1, j, k, ny, nz)];
        if (ic2 && i > 0 && k + 1 < nz) sum = sum - Mtw[P] * z[IDX(i - 1, j, k + 1, ny, nz)];
        if (j > 0) sum = sum - Ms[P] * z[IDX(i, j - 1, k, ny, nz)];
        if (ic2 && j > 0 && k + 1 < nz) sum = sum - Mts[P] * z[IDX(i, j - 1, k + 1, ny, nz)];
        if (k > 0) sum = sum - Mb[P] * z[P - 1];
        z[P] = sum / Mc[P];
      }
  for (int i = nx - 1; i >= 0; i--)
    for (int j = ny - 1; j >= 0; j--)
      for (int k = nz - 1; k >= 0; k--) {
        const size_t P = IDX(i, j, k, ny, nz);
        double sum = z[P];
        if (i + 1 < nx) {
          const size_t E = IDX(i + 1, j, k, ny, nz);
          sum = sum - Mw[E] * z[E];
          if (ic2 && k > 0) sum = sum - Mtw[E - 1] * z[E - 1];
        }
        if (j + 1 < ny) {
          const size_t Nn = IDX(i, j + 1, k, ny, nz);
          sum = sum - Ms[Nn] * z[Nn];
          if (ic2 && k > 0) sum = sum - Mts[Nn - 1] * z[Nn - 1];
        }
        if (k + 1 < nz) sum = sum - Mb[P + 1] * z[P + 1];
        z[P] = sum / Mc[P];
      }
}

void ora_prec_apply(const ora_mat *A, int kind, const double *r, double *z) {
  ora_prec M;
  prec_form(&M, A, kind);
  prec_solve(&M, z, r);
  prec_free(&M);
}

/* ------------------------------------------------------------------------ */
/* CG, Src/Solver/Linear/Krylov/cg.cpp:10-182, with the preconditioner kind  */
/* (Diagonal: preconditioner_form.cpp:4-13, preconditioner_solve.cpp:4-11)  */
/* ------------------------------------------------------------------------ */
int ora_cg_prec(const ora_mat *A, int prec, double *x, const double *b, const double *dV,
                int mini, int maxit, double res_rat, double res_tol, double scale,
                int stalecount, double *hist, ora_result *out) {
  const size_t N = mat_n(A);
  double *p = dalloc(N), *q = dalloc(N), *r = dalloc(N), *z = dalloc(N);
  double *win = stalecount > 0 ? dalloc((size_t)stalecount) : NULL;
  double alfa, beta, rho = 0.0, rho_old = 1.0, pq;
  ora_prec M;
  prec_form(&M, A, prec);
  for (int q2 = 0; q2 < stalecount; q2++) win[q2] = UNREAL;

  ora_resid(A, b, x, r);
  double res = sqrt(ora_dot_voldiv_avg(N, r, r, dV)) / scale;
  double res0 = res;
  if (stalecount > 0) win[stalecount - 1] = res;
  if (hist) hist[0] = res;
  ora_result R = {0, 0, 0, res0, res};
  if (res < res_tol) {
    R.converged = 1; R.reason = 7; R.iters = 0;
    goto done;
  }
  int i;
  for (i = 0; i < maxit; i++) {
    prec_solve(&M, z, r);
    rho = ora_dot(N, r, z);
    if (i == 0) {
      memcpy(p, z, N * sizeof(double));
    } else {
      beta = rho / rho_old;
      for (size_t P = 0; P < N; P++) p[P] = z[P] + beta * p[P];
    }
    ora_spmv(A, p, q);
    pq = ora_dot(N, p, q);
    alfa = rho / pq;
    for (size_t P = 0; P < N; P++) x[P] += alfa * p[P];
    for (size_t P = 0; P < N; P++) r[P] -= alfa * q[P];
    res = sqrt(ora_dot_voldiv_avg(N, r, r, dV)) / scale;
    if (hist) hist[i + 1] = res;
    if (res < res_tol && i >= mini - 1) { R.converged = 1; R.reason = 1; break; }
    if (res < res0 * res_rat && i >= mini - 1) { R.converged = 1; R.reason = 2; break; }
    if (stalecount > 0 && stale_check(win, stalecount, res)) { R.reason = 3; break; }
    rho_old = rho;
  }
  R.iters = (i < maxit) ? i + 1 : maxit + 1; /* printed "i+1" (cg.cpp:178) */
  R.res = res;
done:
  if (out) *out = R;
  free(p); free(q); free(r); free(z); free(win);
  prec_free(&M);
  return R.converged;
}

int ora_cg(const ora_mat *A, double *x, const double *b, const double *dV,
           int mini, int maxit, double res_rat, double res_tol, double scale,
           int stalecount, double *hist, ora_result *out) {
  return ora_cg_prec(A, ORA_PREC_DI, x, b, dV, mini, maxit, res_rat, res_tol, scale,
                     stalecount, hist, out);
}

/* ------------------------------------------------------------------------ */
/* BiCGS, Src/Solver/Linear/Krylov/bicgs.cpp:9-229                           */
/* ------------------------------------------------------------------------ */
/* z = M^-1 r of a BiCGS step: the diagonal (Diagonal::solve,
 * preconditioner_solve.cpp, as r / c) when ac is NULL, else the multigrid
 * cycle on level 0 (the V-cycle: the K-cycle's steps change the operator
 * from one application to the next) */
static int kcycle_mask = 0; /* levels whose cycle is the K-cycle (ora_set_kcycle) */

typedef struct {
  const double *Mc;
  ora_ac *ac;
  int nu1, nu2;
  double omega, cfac;
} bicgs_prec;

static void bicgs_apply(const bicgs_prec *M, size_t N, const double *r, double *z) {
  if (!M->ac) {
    for (size_t P = 0; P < N; P++) z[P] = r[P] * M->Mc[P];
    return;
  }
  const int keep = kcycle_mask;
  kcycle_mask = 0;
  ora_ac_precond(M->ac, 0, r, z, M->nu1, M->nu2, M->omega, M->cfac);
  kcycle_mask = keep;
}

static int bicgs_run(const ora_mat *A, const bicgs_prec *M, double *x, const double *b,
                     const double *dV, int mini, int maxit, double res_rat, double res_tol,
                     double scale, int stalecount, double *hist, ora_result *out) {
  const size_t N = mat_n(A);
  double *p = dalloc(N), *r = dalloc(N), *s = dalloc(N), *t = dalloc(N),
         *v = dalloc(N), *p_ = dalloc(N), *r_ = dalloc(N), *s_ = dalloc(N);
  double *win = stalecount > 0 ? dalloc((size_t)stalecount) : NULL;
  double alfa = 0.0, beta, omega = 1.0, rho, rho_old = 1.0, r_v, ss, ts, tt;
  for (int q2 = 0; q2 < stalecount; q2++) win[q2] = UNREAL;

  ora_resid(A, b, x, r);
  double res = sqrt(ora_dot_voldiv_avg(N, r, r, dV)) / scale;
  double res0 = res;
  if (stalecount > 0) win[stalecount - 1] = res;
  if (hist) hist[0] = res;
  ora_result R = {0, 0, 0, res0, res};
  if (res < res_tol) {
    R.converged = 1; R.reason = 7; R.iters = 0;
    goto done;
  }
  memcpy(r_, r, N * sizeof(double));
  int i;
  for (i = 0; i < maxit; i++) {
    rho = ora_dot(N, r_, r);
    if (sqrt(fabs(rho)) < res_tol) { R.reason = 4; break; }
    if (i == 0) {
      memcpy(p, r, N * sizeof(double));
    } else {
      beta = (rho / rho_old) * (alfa / omega);
      for (size_t P = 0; P < N; P++) p[P] = r[P] + beta * p[P];
      double bo = beta * omega;
      for (size_t P = 0; P < N; P++) p[P] -= bo * v[P];
    }
    bicgs_apply(M, N, p, p_);
    ora_spmv(A, p_, v);
    r_v = ora_dot(N, r_, v);
    alfa = rho / r_v;
    for (size_t P = 0; P < N; P++) s[P] = r[P] - alfa * v[P];
    ss = ora_dot(N, s, s);
    if (sqrt(ss) < res_tol) {
      for (size_t P = 0; P < N; P++) x[P] += alfa * p_[P];
      R.reason = 5;
      break;
    }
    bicgs_apply(M, N, s, s_);
    ora_spmv(A, s_, t);
    ts = ora_dot(N, t, s);
    tt = ora_dot(N, t, t);
    omega = ts / tt;
    for (size_t P = 0; P < N; P++) x[P] += alfa * p_[P];
    for (size_t P = 0; P < N; P++) x[P] += omega * s_[P];
    for (size_t P = 0; P < N; P++) r[P] = s[P] - omega * t[P];
    res = sqrt(ora_dot_voldiv_avg(N, r, r, dV)) / scale;
    if (hist) hist[i + 1] = res;
    if (res < res_tol && i >= mini - 1) { R.converged = 1; R.reason = 1; break; }
    if (res < res0 * res_rat && i >= mini - 1) { R.converged = 1; R.reason = 2; break; }
    if (omega >= -res_tol && omega <= res_tol) { R.reason = 6; break; }
    if (stalecount > 0 && stale_check(win, stalecount, res)) { R.reason = 3; break; }
    rho_old = rho;
  }
  R.iters = (i < maxit) ? i + 1 : maxit + 1;
  R.res = res;
done:
  if (out) *out = R;
  free(p); free(r); free(s); free(t); free(v); free(p_); free(r_); free(s_);
  free(win);
  return R.converged;
}

int ora_bicgs(const ora_mat *A, double *x, const double *b, const double *dV,
              int mini, int maxit, double res_rat, double res_tol, double scale,
              int stalecount, double *hist, ora_result *out) {
  const size_t N = mat_n(A);
  double *Mc = dalloc(N);
  for (size_t P = 0; P < N; P++) Mc[P] = 1.0 / A->c[P];
  const bicgs_prec M = {Mc, NULL, 0, 0, 0.0, 0.0};
  const int ok = bicgs_run(A, &M, x, b, dV, mini, maxit, res_rat, res_tol, scale, stalecount,
                           hist, out);
  free(Mc);
  return ok;
}

/* ------------------------------------------------------------------------ */
/* Jacobi, Src/Solver/Linear/Iterative/jacobi.cpp:10-75 (x = ci (r + c x))   */
/* ------------------------------------------------------------------------ */
int ora_jacobi(const ora_mat *A, double *x, const double *b, int mini, int maxit,
               double res_rat, double res_tol, ora_result *out) {
  const size_t N = mat_n(A);
  double *r = dalloc(N);
  ora_resid(A, b, x, r);
  double res = ora_dot(N, r, r);
  double res0 = res;
  ora_result R = {0, 0, 0, sqrt(res0), sqrt(res)};
  if (sqrt(res) < res_tol) {
    R.converged = 1; R.reason = 7;
    goto done;
  }
  int i;
  for (i = 0; i < maxit; i++) {
    for (size_t P = 0; P < N; P++) {
      double ci = 1.0 / A->c[P];
      x[P] = ci * (r[P] + A->c[P] * x[P]);
    }
    ora_resid(A, b, x, r);
    res = ora_dot(N, r, r);
    if (sqrt(res) < res_tol && i >= mini - 1) { R.converged = 1; R.reason = 1; break; }
    if (sqrt(res) < sqrt(res0) * res_rat && i >= mini - 1) { R.converged = 1; R.reason = 2; break; }
  }
  R.iters = (i < maxit) ? i + 1 : maxit + 1;
  R.res = sqrt(res);
done:
  if (out) *out = R;
  free(r);
  return R.converged;
}

/* ------------------------------------------------------------------------ */
/* AC level construction                                                     */
/* ------------------------------------------------------------------------ */
/* Src/Global/global_factor.cpp:2-19: the first factor found is n/m for the
 * largest proper divisor m, i.e. the smallest prime factor; a prime n gives n */
int ora_smallest_factor(int n) {
  for (int m = n - 1; m >= 2; m--)
    if (n % m == 0) return n / m;
  return n;
}

int ora_levels(int rule, int nx, int ny, int nz, int maxlev, int *lx, int *ly,
               int *lz) {
  int L = 0;
  lx[0] = nx; ly[0] = ny; lz[0] = nz;
  L = 1;
  while (L < maxlev) {
    int n[3] = {lx[L - 1], ly[L - 1], lz[L - 1]};
    int m[3];
    int any = 0;
    if (rule == 3) {
      /* rule 3: halve x and y while they stay >= 8, then z */
      int xy = 0;
      for (int d = 0; d < 2; d++) {
        m[d] = n[d];
        if (n[d] % 2 == 0 && n[d] / 2 >= 8) { m[d] = n[d] / 2; xy = 1; }
      }
      m[2] = n[2];
      if (!xy && n[2] % 2 == 0 && n[2] / 2 >= 8) { m[2] = n[2] / 2; xy = 1; }
      if (!xy) break;
      lx[L] = m[0]; ly[L] = m[1]; lz[L] = m[2];
      L++;
      continue;
    }
    for (int d = 0; d < 3; d++) {
      m[d] = n[d];
      if (rule == 0) {
        /* Src/Domain/domain_coarsen.cpp:33-71 */
        int f = ora_smallest_factor(n[d]);
        if (n[d] % f == 0 && n[d] > 4 && n[d] != f) { m[d] = n[d] / f; any = 1; }
      } else {
        /* rule 1: halve while >= 8; rule 2: halve while >= 4 */
        const int lo = rule == 2 ? 4 : 8;
        if (n[d] % 2 == 0 && n[d] / 2 >= lo) { m[d] = n[d] / 2; any = 1; }
      }
    }
    if (!any) break;
    lx[L] = m[0]; ly[L] = m[1]; lz[L] = m[2];
    L++;
  }
  return L;
}

static void mat_alloc(ora_mat *A, int nx, int ny, int nz) {
  size_t N = (size_t)nx * ny * nz;
  A->nx = nx; A->ny = ny; A->nz = nz;
  A->c = dalloc(N); A->w = dalloc(N); A->e = dalloc(N); A->s = dalloc(N);
  A->n = dalloc(N); A->b = dalloc(N); A->t = dalloc(N);
}

static void mat_free(ora_mat *A) {
  free(A->c); free(A->w); free(A->e); free(A->s); free(A->n); free(A->b);
  free(A->t);
  memset(A, 0, sizeof(*A));
}

int ora_ac_build(ora_ac *ac, int rule, const ora_mat *A0, const double *dV0,
                 const uint8_t *off0, int maxlev) {
  memset(ac, 0, sizeof(*ac));
  if (maxlev > 24) maxlev = 24;
  ac->nlev = ora_levels(rule, A0->nx, A0->ny, A0->nz, maxlev, ac->nx, ac->ny, ac->nz);
  for (int l = 0; l < ac->nlev; l++) {
    size_t N = (size_t)ac->nx[l] * ac->ny[l] * ac->nz[l];
    mat_alloc(&ac->A[l], ac->nx[l], ac->ny[l], ac->nz[l]);
    ac->aflag[l] = (int *)calloc(N, sizeof(int));
    ac->dV[l] = dalloc(N);
    ac->phi[l] = dalloc(N);
    ac->f[l] = dalloc(N);
    ac->res[l] = dalloc(N);
  }
  size_t N0 = (size_t)A0->nx * A0->ny * A0->nz;
  memcpy(ac->A[0].c, A0->c, N0 * 8); memcpy(ac->A[0].w, A0->w, N0 * 8);
  memcpy(ac->A[0].e, A0->e, N0 * 8); memcpy(ac->A[0].s, A0->s, N0 * 8);
  memcpy(ac->A[0].n, A0->n, N0 * 8); memcpy(ac->A[0].b, A0->b, N0 * 8);
  memcpy(ac->A[0].t, A0->t, N0 * 8);
  memcpy(ac->dV[0], dV0, N0 * 8);
  for (size_t P = 0; P < N0; P++) ac->aflag[0][P] = off0 ? !off0[P] : 1;
  for (int l = 1; l < ac->nlev; l++) {
    ora_coarsen_flag(ac, l);
    ora_coarsen_system(ac, l);
    /* coarse cell volume: sum over the block (the reference's coarse grid
     * spans the same nodes, so this equals its dV up to rounding) */
    int CI = ac->nx[l - 1] / ac->nx[l], CJ = ac->ny[l - 1] / ac->ny[l],
        CK = ac->nz[l - 1] / ac->nz[l];
    for (int I = 0; I < ac->nx[l]; I++)
      for (int J = 0; J < ac->ny[l]; J++)
        for (int K = 0; K < ac->nz[l]; K++) {
          double s = 0.0;
          for (int a = 0; a < CI; a++)
            for (int bb = 0; bb < CJ; bb++)
              for (int cc = 0; cc < CK; cc++)
                s += ac->dV[l - 1][IDX(I * CI + a, J * CJ + bb, K * CK + cc,
                                        ac->ny[l - 1], ac->nz[l - 1])];
          ac->dV[l][IDX(I, J, K, ac->ny[l], ac->nz[l])] = s;
        }
  }
  return ac->nlev;
}

void ora_ac_free(ora_ac *ac) {
  for (int l = 0; l < ac->nlev; l++) {
    mat_free(&ac->A[l]);
    free(ac->aflag[l]); free(ac->dV[l]); free(ac->phi[l]); free(ac->f[l]);
    free(ac->res[l]);
  }
  free(ac->G);
  memset(ac, 0, sizeof(*ac));
}

/* Src/Solver/Additive/additive_coarsen_flag.cpp:4-41: OR over the block */
void ora_coarsen_flag(const ora_ac *ac, int l) {
  const int nX = ac->nx[l], nY = ac->ny[l], nZ = ac->nz[l];
  const int nx = ac->nx[l - 1], ny = ac->ny[l - 1], nz = ac->nz[l - 1];
  const int CI = nx / nX, CJ = ny / nY, CK = nz / nZ;
  for (int I = 0; I < nX; I++)
    for (int J = 0; J < nY; J++)
      for (int K = 0; K < nZ; K++) {
        int fl = 0;
        for (int a = 0; a < CI; a++)
          for (int bb = 0; bb < CJ; bb++)
            for (int cc = 0; cc < CK; cc++)
              fl |= ac->aflag[l - 1][IDX(I * CI + a, J * CJ + bb, K * CK + cc, ny, nz)];
        ac->aflag[l][IDX(I, J, K, nY, nZ)] = fl;
      }
}

/* Src/Solver/Additive/additive_coarsen_system.cpp:4-140.  Within a block the
 * fine cells are visited as the reference does: c summed over (i,j,k), then
 * the internal e/w couplings per (j,k) line, then n/s per (i,k), then t/b
 * per (i,j); the external couplings are the sums over the block's faces. */
void ora_coarsen_system(const ora_ac *ac, int l) {
  const int nX = ac->nx[l], nY = ac->ny[l], nZ = ac->nz[l];
  const int ny = ac->ny[l - 1], nz = ac->nz[l - 1];
  const int CI = ac->nx[l - 1] / nX, CJ = ny / nY, CK = nz / nZ;
  const ora_mat *h = &ac->A[l - 1];
  const ora_mat *H = &ac->A[l];
#define F(i, j, k) IDX(I * CI + (i), J * CJ + (j), K * CK + (k), ny, nz)
  for (int I = 0; I < nX; I++)
    for (int J = 0; J < nY; J++)
      for (int K = 0; K < nZ; K++) {
        size_t Q = IDX(I, J, K, nY, nZ);
        double c = 0.0;
        for (int i = 0; i < CI; i++)
          for (int j = 0; j < CJ; j++)
            for (int k = 0; k < CK; k++) c += h->c[F(i, j, k)];
        for (int j = 0; j < CJ; j++)
          for (int k = 0; k < CK; k++) {
            for (int i = 0; i < CI - 1; i++) c -= h->e[F(i, j, k)];
            for (int i = 1; i < CI; i++) c -= h->w[F(i, j, k)];
          }
        double e = 0.0, w = 0.0;
        for (int j = 0; j < CJ; j++)
          for (int k = 0; k < CK; k++) {
            e += h->e[F(CI - 1, j, k)];
            w += h->w[F(0, j, k)];
          }
        for (int i = 0; i < CI; i++)
          for (int k = 0; k < CK; k++) {
            for (int j = 0; j < CJ - 1; j++) c -= h->n[F(i, j, k)];
            for (int j = 1; j < CJ; j++) c -= h->s[F(i, j, k)];
          }
        double n = 0.0, s = 0.0;
        for (int i = 0; i < CI; i++)
          for (int k = 0; k < CK; k++) {
            n += h->n[F(i, CJ - 1, k)];
            s += h->s[F(i, 0, k)];
          }
        for (int i = 0; i < CI; i++)
          for (int j = 0; j < CJ; j++) {
            for (int k = 0; k < CK - 1; k++) c -= h->t[F(i, j, k)];
            for (int k = 1; k < CK; k++) c -= h->b[F(i, j, k)];
          }
        double t = 0.0, b = 0.0;
        for (int i = 0; i < CI; i++)
          for (int j = 0; j < CJ; j++) {
            t += h->t[F(i, j, CK - 1)];
            b += h->b[F(i, j, 0)];
          }
        H->c[Q] = c; H->e[Q] = e; H->w[Q] = w; H->n[Q] = n; H->s[Q] = s;
        H->t[Q] = t; H->b[Q] = b;
      }
#undef F
}

/* Src/Solver/Additive/additive_restriction.cpp:12-177: H.f = sum over the
 * block of (f - A phi)_h, written as sum f + external couplings * phi outside
 * - D phi with D = c - internal couplings; inactive cells 0; mean over active
 * cells removed. */
void ora_restriction(const ora_ac *ac, int l) {
  const int nx = ac->nx[l], ny = ac->ny[l], nz = ac->nz[l];
  const int nX = ac->nx[l + 1], nY = ac->ny[l + 1], nZ = ac->nz[l + 1];
  const int CI = nx / nX, CJ = ny / nY, CK = nz / nZ;
  const ora_mat *h = &ac->A[l];
  const double *phi = ac->phi[l], *f = ac->f[l];
  double *fH = ac->f[l + 1];
  double *D = dalloc((size_t)CI * CJ * CK);
#define DD(i, j, k) D[((size_t)(i) * CJ + (j)) * CK + (k)]
  for (int I = 0; I < nX; I++)
    for (int J = 0; J < nY; J++)
      for (int K = 0; K < nZ; K++) {
        const int i0 = I * CI, j0 = J * CJ, k0 = K * CK;
#define F(i, j, k) IDX(i0 + (i), j0 + (j), k0 + (k), ny, nz)
#define PH(i, j, k) nbx(phi, i0 + (i), j0 + (j), k0 + (k), nx, ny, nz)
        double s = 0.0;
        for (int i = 0; i < CI; i++)
          for (int j = 0; j < CJ; j++)
            for (int k = 0; k < CK; k++) {
              DD(i, j, k) = h->c[F(i, j, k)];
              s += f[F(i, j, k)];
            }
        for (int j = 0; j < CJ; j++)
          for (int k = 0; k < CK; k++) {
            s += h->e[F(CI - 1, j, k)] * PH(CI, j, k);
            s += h->w[F(0, j, k)] * PH(-1, j, k);
            for (int i = 0; i < CI - 1; i++) DD(i, j, k) -= h->e[F(i, j, k)];
            for (int i = CI - 1; i > 0; i--) DD(i, j, k) -= h->w[F(i, j, k)];
          }
        for (int i = 0; i < CI; i++)
          for (int k = 0; k < CK; k++) {
            s += h->n[F(i, CJ - 1, k)] * PH(i, CJ, k);
            s += h->s[F(i, 0, k)] * PH(i, -1, k);
            for (int j = 0; j < CJ - 1; j++) DD(i, j, k) -= h->n[F(i, j, k)];
            for (int j = CJ - 1; j > 0; j--) DD(i, j, k) -= h->s[F(i, j, k)];
          }
        for (int i = 0; i < CI; i++)
          for (int j = 0; j < CJ; j++) {
            s += h->t[F(i, j, CK - 1)] * PH(i, j, CK);
            s += h->b[F(i, j, 0)] * PH(i, j, -1);
            for (int k = 0; k < CK - 1; k++) DD(i, j, k) -= h->t[F(i, j, k)];
            for (int k = CK - 1; k > 0; k--) DD(i, j, k) -= h->b[F(i, j, k)];
          }
        for (int i = 0; i < CI; i++)
          for (int j = 0; j < CJ; j++)
            for (int k = 0; k < CK; k++) s -= DD(i, j, k) * phi[F(i, j, k)];
        fH[IDX(I, J, K, nY, nZ)] = s;
#undef F
#undef PH
      }
#undef DD
  free(D);
  const size_t NH = (size_t)nX * nY * nZ;
  const int *fl = ac->aflag[l + 1];
  for (size_t Q = 0; Q < NH; Q++)
    if (!fl[Q]) fH[Q] = 0.0;
  double sum = 0.0;
  long ncell = 0;
  for (size_t Q = 0; Q < NH; Q++)
    if (fl[Q]) { sum += fH[Q]; ncell++; }
  sum /= (double)ncell;
  for (size_t Q = 0; Q < NH; Q++)
    if (fl[Q]) fH[Q] -= sum;
}

/* Src/Solver/Additive/additive_interpolation.cpp:12-64: h.phi += H.phi */
void ora_interpolation(const ora_ac *ac, int l) {
  const int nX = ac->nx[l], nY = ac->ny[l], nZ = ac->nz[l];
  const int ny = ac->ny[l - 1], nz = ac->nz[l - 1];
  const int CI = ac->nx[l - 1] / nX, CJ = ny / nY, CK = nz / nZ;
  for (int I = 0; I < nX; I++)
    for (int J = 0; J < nY; J++)
      for (int K = 0; K < nZ; K++) {
        double v = ac->phi[l][IDX(I, J, K, nY, nZ)];
        for (int i = 0; i < CI; i++)
          for (int j = 0; j < CJ; j++)
            for (int k = 0; k < CK; k++)
              ac->phi[l - 1][IDX(I * CI + i, J * CJ + j, K * CK + k, ny, nz)] += v;
      }
}

/* AC::residual (additive_residual.cpp:19-29): sqrt(dot_voldiv_avg) * dt,
 * and (linf != NULL) max |r / dV| * dt (Scalar::max_abs_voldiv,
 * scalar_minmax.cpp:155-165) */
static double ac_residual2(ora_ac *ac, double dt, double *linf) {
  const ora_mat *A = &ac->A[0];
  size_t N = mat_n(A);
  ora_resid(A, ac->f[0], ac->phi[0], ac->res[0]);
  const double r2 = sqrt(ora_dot_voldiv_avg(N, ac->res[0], ac->res[0], ac->dV[0])) * dt;
  if (linf) {
    double vmax = fabs(ac->res[0][0] / ac->dV[0][0]);
    for (size_t P = 0; P < N; P++)
      if (fabs(ac->res[0][P] / ac->dV[0][P]) > vmax) vmax = fabs(ac->res[0][P] / ac->dV[0][P]);
    *linf = vmax * dt;
  }
  return r2;
}
static double ac_residual(ora_ac *ac, double dt) { return ac_residual2(ac, dt, NULL); }

/* the recursive cycles of additive_recursive.cpp:4-166 with AC's level
 * tolerances resrat / restol (atto by default, additive.cpp:16-17) and the
 * coarsest call MaxIter 100, ResRat -1 */
typedef struct {
  const int *mv, *ms;  /* MaxIter and stale counts: pre / re- / post-smoothing */
  double rr, rt, scale;
  int solver;          /* 0 CG, 1 BiCGS */
} cyc_par;

/* Linear::solve of one level (call_smoother / call_solver,
 * additive_cycle_functions.cpp:90-107): the equation's Krylov solver, CG or
 * BiCGS with the diagonal preconditioner; returns whether it met its
 * tolerance */
static int ac_cg(ora_ac *ac, int l, int mi, double rr, double rt, double scale, int ms,
                 int solver) {
  if (solver == 1)
    return ora_bicgs(&ac->A[l], ac->phi[l], ac->f[l], ac->dV[l], 1, mi, rr, rt, scale, ms,
                     NULL, NULL);
  return ora_cg(&ac->A[l], ac->phi[l], ac->f[l], ac->dV[l], 1, mi, rr, rt, scale, ms,
                NULL, NULL);
}

static void ac_coarsest(ora_ac *ac, int l, const cyc_par *q) {
  ac_cg(ac, l, 100, -1.0, q->rt, q->scale, -1, q->solver);
  ora_interpolation(ac, l);
}

static void ac_v(ora_ac *ac, int l, const cyc_par *q) {
  if (l != ac->nlev - 1) {
    ac_cg(ac, l, q->mv[0], q->rr, q->rt, q->scale, q->ms[0], q->solver);
    ora_restriction(ac, l);
    ac_v(ac, l + 1, q);
    ac_cg(ac, l, q->mv[2], q->rr, q->rt, q->scale, q->ms[2], q->solver);
    if (l > 0) ora_interpolation(ac, l);
  } else {
    ac_coarsest(ac, l, q);
  }
}

/* f_cycle (:37-79) and w_cycle (:82-124): smooth, restrict, recurse (F: an
 * F-cycle, W: a W-cycle), re-smooth with mv[1], restrict again, recurse on
 * the way back (F: a V-cycle, W: a W-cycle) with the coarse phi kept from
 * the first descent, post-smooth, prolongate */
static void ac_fw(ora_ac *ac, int l, const cyc_par *q, int w) {
  if (l != ac->nlev - 1) {
    ac_cg(ac, l, q->mv[0], q->rr, q->rt, q->scale, q->ms[0], q->solver);
    ora_restriction(ac, l);
    ac_fw(ac, l + 1, q, w);
    ac_cg(ac, l, q->mv[1], q->rr, q->rt, q->scale, q->ms[1], q->solver);
    ora_restriction(ac, l);
    if (w) ac_fw(ac, l + 1, q, 1);
    else ac_v(ac, l + 1, q);
    ac_cg(ac, l, q->mv[2], q->rr, q->rt, q->scale, q->ms[2], q->solver);
    if (l > 0) ora_interpolation(ac, l);
  } else {
    ac_coarsest(ac, l, q);
  }
}

static void ac_loop(ora_ac *ac, int l, int kind, const cyc_par *q) {
  if (kind == ORA_CYC_V) ac_v(ac, l, q);
  else ac_fw(ac, l, q, kind == ORA_CYC_W);
}

/* flex_cycle (additive_recursive.cpp:161-200) from level 0, its tail
 * recursion as a loop: smooth (MaxIter mv[0] on the way down, mv[2] on the
 * way up); a level whose smoother met its tolerance hands its phi up
 * (interpolation) and the finer level smooths again, level 0 ends the cycle;
 * a level whose smoother did not restricts its residual and goes down; the
 * coarsest level is solved (MaxIter 100) and hands up.  The reference bounds
 * nothing (once a level's smoother cannot meet its ResRat the cycle
 * alternates between two levels for good); here a cycle ends after
 * max_visits level visits: returns 1 then, 0 when level 0 ended it. */
static int ac_flex(ora_ac *ac, const cyc_par *q, int max_visits) {
  int l = 0, up = 0;
  for (int v = 0; v < max_visits; v++) {
    if (l != ac->nlev - 1) {
      const int s = up ? 2 : 0;
      if (ac_cg(ac, l, q->mv[s], q->rr, q->rt, q->scale, q->ms[s], q->solver)) {
        if (l == 0) return 0;
        ora_interpolation(ac, l);
        l--;
        up = 1;
      } else {
        ora_restriction(ac, l);
        l++;
        up = 0;
      }
    } else {
      ac_coarsest(ac, l, q);
      l--;
      up = 0;  /* the default sign of the call (additive.h:185-187) */
    }
  }
  return 1;
}

/* full_cycle (:127-163): restrict, recurse to the coarsest level, then run
 * the chosen cycle from each level on the way back */
static void ac_full(ora_ac *ac, int l, int kind, const cyc_par *q) {
  if (l != ac->nlev - 1) {
    ora_restriction(ac, l);
    ac_full(ac, l + 1, kind, q);
    ac_loop(ac, l, kind, q);
  } else {
    ac_coarsest(ac, l, q);
  }
}

void ora_cycle_defaults(ora_cycle *P) {
  P->toler = -1.0;
  P->factor = -1.0;
  P->max_cyc = 20;
  P->min_cyc = 0;
  P->mv[0] = P->mv[1] = P->mv[2] = 20;
  P->ms[0] = P->ms[1] = P->ms[2] = -1;
  P->dt = 1.0;
  P->init = ORA_CYC_NONE;
  P->loop = ORA_CYC_V;
  P->stop_if_div = 1;
  P->priority_min_cyc = 1;
  P->use_linf = 0;
  P->resrat = ATTO;
  P->restol = ATTO;
  P->max_visits = 1000;
  P->solver = 0;
}

/* AC::converged (additive_cycle_functions.cpp:49-87) */
static int ac_converged(ora_ac *ac, const ora_cycle *P, int cycle, double *res1,
                        double *linf1, double res_beg, double linf_beg, double res0,
                        double linf0, const double *fold) {
  *res1 = ac_residual2(ac, P->dt, linf1);
  const double rc = P->use_linf ? *linf1 : *res1;
  const double r0c = P->use_linf ? linf0 : res0;
  const double rbc = P->use_linf ? linf_beg : res_beg;
  if (cycle >= P->min_cyc && (rc / rbc < P->factor || rc < P->toler)) return 0;
  int check = P->stop_if_div;
  if (P->priority_min_cyc && cycle < P->min_cyc) check = 0;
  if (check && rc >= r0c) {
    memcpy(ac->phi[0], fold, mat_n(&ac->A[0]) * 8);
    return 1;
  }
  return 2;
}

/* AC::cycle (additive_cycle.cpp:4-99) with init_cycles
 * (additive_cycle_functions.cpp:12-46).  hist / hist_linf (max_cyc + 2
 * entries): [0] the initial residual, then one per evaluation (the
 * full-cycle initialisation's first when init is set). */
int ora_ac_cycle(ora_ac *ac, double *phi0, const double *f0, const ora_cycle *P,
                 double *hist, double *hist_linf, int *nhist, int *ncyc) {
  const size_t N0 = mat_n(&ac->A[0]);
  const cyc_par q = {P->mv, P->ms, P->resrat, P->restol, 1.0 / P->dt * 1.0, P->solver};
  double *fold = dalloc(N0);
  int ret = 0, nh = 0;
  memcpy(ac->phi[0], phi0, N0 * 8);
  memcpy(ac->f[0], f0, N0 * 8);
  double linf_0 = 0.0;
  const double res_0 = ac_residual2(ac, P->dt, &linf_0);
#define PUSH(a, b) do { if (hist) hist[nh] = (a); if (hist_linf) hist_linf[nh] = (b); nh++; } while (0)
  PUSH(res_0, linf_0);
  memcpy(fold, ac->phi[0], N0 * 8);
  const double r0c = P->use_linf ? linf_0 : res_0;
  if ((r0c < P->toler && P->min_cyc == 0) || res_0 == 0.0) {
    if (ncyc) *ncyc = 0;
    ret = 1;
    goto done;
  }
  for (int l = 1; l < ac->nlev; l++) {
    size_t N = (size_t)ac->nx[l] * ac->ny[l] * ac->nz[l];
    memset(ac->phi[l], 0, N * 8);
    memset(ac->f[l], 0, N * 8);
  }
  double res_m0 = 0.0, linf_m0 = 0.0, res_m1, linf_m1;
  if (P->init != ORA_CYC_NONE) {
    ac_full(ac, 0, P->init, &q);
    res_m1 = ac_residual2(ac, P->dt, &linf_m1);
    /* the verdict is not used (additive_cycle.cpp:40-42): with the
     * divergence check on, the equal residuals restore fold */
    (void)ac_converged(ac, P, 0, &res_m0, &linf_m0, res_0, linf_0, res_m1, linf_m1, fold);
    PUSH(res_m0, linf_m0);
  } else {
    res_m1 = ac_residual2(ac, P->dt, &linf_m1);
  }
  if (ncyc) *ncyc = P->max_cyc;
  for (int c = 1; c <= P->max_cyc; c++) {
    if (P->loop == ORA_CYC_FLEX) {
      if (ac_flex(ac, &q, P->max_visits)) {
        /* no end in sight: the last cycle's phi, as after a divergence */
        memcpy(ac->phi[0], fold, N0 * 8);
        if (ncyc) *ncyc = c;
        ret = -1;
        goto done;
      }
    } else {
      ac_loop(ac, 0, P->loop, &q);
    }
    const int con = ac_converged(ac, P, c, &res_m0, &linf_m0, res_0, linf_0, res_m1, linf_m1,
                                 fold);
    PUSH(res_m0, linf_m0);
    res_m1 = res_m0;  /* linf_m1 keeps its value (additive_cycle.cpp:78) */
    if (con == 0 || con == 1) {
      if (ncyc) *ncyc = c;
      ret = con == 0;
      goto done;
    }
    memcpy(fold, ac->phi[0], N0 * 8);
  }
  ret = 0;
#undef PUSH
done:
  if (nhist) *nhist = nh;
  memcpy(phi0, ac->phi[0], N0 * 8);
  free(fold);
  return ret;
}

/* AC::vcycle: the V loop with the constructor's controls */
int ora_ac_vcycle(ora_ac *ac, double *phi0, const double *f0, double toler,
                  double factor, int max_cyc, int min_cyc, const int mv[3],
                  double dt, double *hist, int *ncyc) {
  ora_cycle P;
  ora_cycle_defaults(&P);
  P.toler = toler;
  P.factor = factor;
  P.max_cyc = max_cyc;
  P.min_cyc = min_cyc;
  for (int k = 0; k < 3; k++) P.mv[k] = mv[k];
  P.dt = dt;
  return ora_ac_cycle(ac, phi0, f0, &P, hist, NULL, NULL, ncyc);
}

/* ------------------------------------------------------------------------ */
/* benchmark composition: V-cycle preconditioned CG (not in the reference)   */
/* ------------------------------------------------------------------------ */
/* e, n, t equal the neighbours' w, s, b (inside the box) */
static int mat_symmetric(const ora_mat *A) {
  const int nx = A->nx, ny = A->ny, nz = A->nz;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        size_t P = IDX(i, j, k, ny, nz);
        if (i < nx - 1 && A->e[P] != A->w[IDX(i + 1, j, k, ny, nz)]) return 0;
        if (j < ny - 1 && A->n[P] != A->s[IDX(i, j + 1, k, ny, nz)]) return 0;
        if (k < nz - 1 && A->t[P] != A->b[IDX(i, j, k + 1, ny, nz)]) return 0;
      }
  return 1;
}

/* G = A^-1 on the active cells (0 elsewhere) of a non-symmetric coarsest
 * level, by LU with partial pivoting (Golub and Van Loan, Matrix
 * Computations, alg. 3.4.1) and a solve per unit column; such a level carries
 * a time term: a pivot below 1e-14 of the largest |c| is reported (-1) */
static int coarse_inverse(ora_ac *ac) {
  const int l = ac->nlev - 1;
  const ora_mat *A = &ac->A[l];
  const int nx = A->nx, ny = A->ny, nz = A->nz;
  const int m = nx * ny * nz;
  const int *fl = ac->aflag[l];
  int *pos = (int *)malloc(sizeof(int) * (size_t)m), *idx = (int *)malloc(sizeof(int) * (size_t)m);
  int na = 0;
  for (int P = 0; P < m; P++) {
    pos[P] = fl[P] ? na : -1;
    if (fl[P]) idx[na++] = P;
  }
  double *U = dalloc((size_t)na * na), *y = dalloc((size_t)na), cmax = 0.0;
  int *perm = (int *)malloc(sizeof(int) * (size_t)(na ? na : 1));
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        const int P = (int)IDX(i, j, k, ny, nz);
        if (!fl[P]) continue;
        double *row = U + (size_t)pos[P] * na;
        row[pos[P]] = A->c[P];
        if (fabs(A->c[P]) > cmax) cmax = fabs(A->c[P]);
        int Q;
        if (i > 0 && fl[Q = (int)IDX(i - 1, j, k, ny, nz)]) row[pos[Q]] = -A->w[P];
        if (i < nx - 1 && fl[Q = (int)IDX(i + 1, j, k, ny, nz)]) row[pos[Q]] = -A->e[P];
        if (j > 0 && fl[Q = (int)IDX(i, j - 1, k, ny, nz)]) row[pos[Q]] = -A->s[P];
        if (j < ny - 1 && fl[Q = (int)IDX(i, j + 1, k, ny, nz)]) row[pos[Q]] = -A->n[P];
        if (k > 0 && fl[Q = (int)IDX(i, j, k - 1, ny, nz)]) row[pos[Q]] = -A->b[P];
        if (k < nz - 1 && fl[Q = (int)IDX(i, j, k + 1, ny, nz)]) row[pos[Q]] = -A->t[P];
      }
  int ok = 1;
  for (int q = 0; q < na; q++) perm[q] = q;
  for (int k = 0; k < na && ok; k++) {
    int p = k;
    for (int i = k + 1; i < na; i++)
      if (fabs(U[(size_t)i * na + k]) > fabs(U[(size_t)p * na + k])) p = i;
    if (!(fabs(U[(size_t)p * na + k]) > 1e-14 * cmax)) { ok = 0; break; }
    if (p != k) {
      for (int j = 0; j < na; j++) {
        double tmp = U[(size_t)p * na + j];
        U[(size_t)p * na + j] = U[(size_t)k * na + j];
        U[(size_t)k * na + j] = tmp;
      }
      int tp = perm[p]; perm[p] = perm[k]; perm[k] = tp;
    }
    for (int i = k + 1; i < na; i++) {
      const double lik = U[(size_t)i * na + k] / U[(size_t)k * na + k];
      U[(size_t)i * na + k] = lik;
      for (int j = k + 1; j < na; j++) U[(size_t)i * na + j] -= lik * U[(size_t)k * na + j];
    }
  }
  if (ok) {
    free(ac->G);
    ac->G = dalloc((size_t)m * m);
    for (int col = 0; col < na; col++) {
      for (int i = 0; i < na; i++) y[i] = perm[i] == col ? 1.0 : 0.0;
      for (int i = 0; i < na; i++)
        for (int k = 0; k < i; k++) y[i] -= U[(size_t)i * na + k] * y[k];
      for (int i = na - 1; i >= 0; i--) {
        for (int k = i + 1; k < na; k++) y[i] -= U[(size_t)i * na + k] * y[k];
        y[i] /= U[(size_t)i * na + i];
      }
      for (int i = 0; i < na; i++) ac->G[(size_t)idx[i] * m + idx[col]] = y[i];
    }
  }
  free(pos); free(idx); free(U); free(y); free(perm);
  return ok ? m : -1;
}

int ora_ac_coarse_dense(ora_ac *ac) {
  const int l = ac->nlev - 1;
  const ora_mat *A = &ac->A[l];
  if (!mat_symmetric(A)) return coarse_inverse(ac);
  const int nx = A->nx, ny = A->ny, nz = A->nz;
  const int m = nx * ny * nz;
  double *M = dalloc((size_t)m * m);
  double *B = dalloc((size_t)m * m);
  int *pin = (int *)calloc((size_t)m, sizeof(int));
  const int *fl = ac->aflag[l];
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        int P = (int)IDX(i, j, k, ny, nz);
        if (!fl[P]) continue;
        M[(size_t)P * m + P] = A->c[P];
        if (i > 0) M[(size_t)P * m + IDX(i - 1, j, k, ny, nz)] = -A->w[P];
        if (i < nx - 1) M[(size_t)P * m + IDX(i + 1, j, k, ny, nz)] = -A->e[P];
        if (j > 0) M[(size_t)P * m + IDX(i, j - 1, k, ny, nz)] = -A->s[P];
        if (j < ny - 1) M[(size_t)P * m + IDX(i, j + 1, k, ny, nz)] = -A->n[P];
        if (k > 0) M[(size_t)P * m + IDX(i, j, k - 1, ny, nz)] = -A->b[P];
        if (k < nz - 1) M[(size_t)P * m + IDX(i, j, k + 1, ny, nz)] = -A->t[P];
      }
  /* Cholesky (lower, in place) with pinning of vanishing pivots: a pivot
   * below 1e-10 of its original diagonal is the null direction of a
   * connected Neumann block; that unknown is fixed to zero. */
  for (int P = 0; P < m; P++)
    if (!fl[P]) pin[P] = 1;
  for (int j = 0; j < m; j++) {
    if (pin[j]) continue;
    double d = M[(size_t)j * m + j];
    double orig = d;
    for (int k = 0; k < j; k++)
      if (!pin[k]) d -= M[(size_t)j * m + k] * M[(size_t)j * m + k];
    (void)orig;
    if (d <= 1e-10 * fabs(A->c[j])) { pin[j] = 1; continue; }
    double Ljj = sqrt(d);
    M[(size_t)j * m + j] = Ljj;
    for (int i = j + 1; i < m; i++) {
      if (pin[i]) continue;
      double s = M[(size_t)i * m + j];
      for (int k = 0; k < j; k++)
        if (!pin[k]) s -= M[(size_t)i * m + k] * M[(size_t)j * m + k];
      M[(size_t)i * m + j] = s / Ljj;
    }
  }
  /* B = (L L^T)^-1 on unpinned rows/cols, zero elsewhere */
  double *y = dalloc((size_t)m);
  for (int col = 0; col < m; col++) {
    if (pin[col]) continue;
    for (int i = 0; i < m; i++) y[i] = (i == col) ? 1.0 : 0.0;
    for (int i = 0; i < m; i++) {
      if (pin[i]) { y[i] = 0.0; continue; }
      double s = y[i];
      for (int k = 0; k < i; k++)
        if (!pin[k]) s -= M[(size_t)i * m + k] * y[k];
      y[i] = s / M[(size_t)i * m + i];
    }
    for (int i = m - 1; i >= 0; i--) {
      if (pin[i]) { y[i] = 0.0; continue; }
      double s = y[i];
      for (int k = i + 1; k < m; k++)
        if (!pin[k]) s -= M[(size_t)k * m + i] * y[k];
      y[i] = s / M[(size_t)i * m + i];
    }
    for (int i = 0; i < m; i++) B[(size_t)i * m + col] = y[i];
  }
  /* G = Pi B Pi over active cells; B itself for a non-singular operator:
   * an active row of the finest matrix that does not sum to zero (Dirichlet
   * sides).  The finest level decides: on the coarse levels the Galerkin
   * sums' cancellation and the scaled solid cells leave row sums of a
   * singular operator well above rounding. */
  int na = 0, singular = 1;
  for (int P = 0; P < m; P++) na += fl[P] ? 1 : 0;
  {
    const ora_mat *F = &ac->A[0];
    const size_t NF = mat_n(F);
    for (size_t P = 0; P < NF && singular; P++) {
      if (!ac->aflag[0][P]) continue;
      const double off = F->w[P] + F->e[P] + F->s[P] + F->n[P] + F->b[P] + F->t[P];
      if (fabs(F->c[P] - off) > 1e-10 * fabs(F->c[P])) singular = 0;
    }
  }
  if (!singular) {
    free(ac->G);
    ac->G = B;
    free(y); free(M); free(pin);
    return m;
  }
  double *rm = dalloc((size_t)m), *cm = dalloc((size_t)m);
  double tm = 0.0;
  for (int i = 0; i < m; i++) {
    if (!fl[i]) continue;
    for (int j = 0; j < m; j++) {
      if (!fl[j]) continue;
      rm[i] += B[(size_t)i * m + j];
      cm[j] += B[(size_t)i * m + j];
      tm += B[(size_t)i * m + j];
    }
  }
  for (int i = 0; i < m; i++) { rm[i] /= na; cm[i] /= na; }
  tm /= (double)na * (double)na;
  free(ac->G);
  ac->G = dalloc((size_t)m * m);
  for (int i = 0; i < m; i++)
    for (int j = 0; j < m; j++)
      ac->G[(size_t)i * m + j] =
          (fl[i] && fl[j]) ? B[(size_t)i * m + j] - rm[i] - cm[j] + tm : 0.0;
  free(rm); free(cm); free(y); free(M); free(B); free(pin);
  return m;
}

static void jacobi_sweep(const ora_mat *A, const double *r, double *e, double *t,
                         double omega) {
  size_t N = mat_n(A);
  ora_resid(A, r, e, t);
  for (size_t P = 0; P < N; P++) e[P] += omega * t[P] / A->c[P];
}

/* z-line Jacobi: e += omega T^-1 (r - A e), T the tridiagonal part of A
 * along k (diagonal c, couplings -b) of each (i, j) column, solved by the
 * Thomas algorithm (forward elimination, back substitution). */
static void zline_solve(const ora_mat *A, const double *t, double *d, double *y) {
  const int nx = A->nx, ny = A->ny, nz = A->nz;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      const size_t P0 = IDX(i, j, 0, ny, nz);
      double piv = A->c[P0];
      y[0] = t[P0];
      d[P0] = piv;
      for (int k = 1; k < nz; k++) {
        const double a = A->b[P0 + k];
        const double l = a / piv;
        piv = A->c[P0 + k] - l * a;
        d[P0 + k] = piv;
        y[k] = t[P0 + k] + l * y[k - 1];
      }
      /* back substitution, solution into d */
      double xn = y[nz - 1] / d[P0 + nz - 1];
      d[P0 + nz - 1] = xn;
      for (int k = nz - 2; k >= 0; k--) {
        xn = (y[k] + A->b[P0 + k + 1] * xn) / d[P0 + k];
        d[P0 + k] = xn;
      }
    }
}

void ora_zline_solve(const ora_mat *A, const double *t, double *x) {
  double *y = dalloc((size_t)A->nz);
  zline_solve(A, t, x, y);
  free(y);
}

static void zline_sweep(const ora_mat *A, const double *r, double *e, double *t,
                        double *d, double *y, double omega) {
  size_t N = mat_n(A);
  ora_resid(A, r, e, t);
  zline_solve(A, t, d, y);
  for (size_t P = 0; P < N; P++) e[P] += omega * d[P];
}

/* levels whose coarse correction is two steps of flexible CG preconditioned
 * by the cycle below (the K-cycle of Notay and Vassilevski, Numer. Linear
 * Algebra Appl. 15 (2008) 473-487): bit l set for level l */
void ora_set_kcycle(int mask) { kcycle_mask = mask; }

/* out ~ A_l^-1 r: c = B r, v = A c, alpha1 = c.r / c.v, rt = r - alpha1 v,
 * d = B rt, gamma = d.v, beta = d.A d, alpha2 = d.rt / (beta - gamma^2 / c.v),
 * out = (alpha1 - gamma alpha2 / c.v) c + alpha2 d, with B the cycle on
 * level l (the second step is always taken) */
static void ac_kstep(ora_ac *ac, int l, const double *r, double *out, int nu1, int nu2,
                     double omega, double cfac, int smoother) {
  const ora_mat *A = &ac->A[l];
  const size_t N = mat_n(A);
  double *c = dalloc(N), *v = dalloc(N), *d = dalloc(N), *w = dalloc(N), *rt = dalloc(N);
  ora_ac_precond_s(ac, l, r, c, nu1, nu2, omega, cfac, smoother);
  ora_spmv(A, c, v);
  const double rho1 = ora_dot(N, c, v);
  const double a1 = rho1 > 0.0 ? ora_dot(N, c, r) / rho1 : 0.0;
  for (size_t P = 0; P < N; P++) rt[P] = r[P] - a1 * v[P];
  ora_ac_precond_s(ac, l, rt, d, nu1, nu2, omega, cfac, smoother);
  ora_spmv(A, d, w);
  const double gam = ora_dot(N, d, v), bet = ora_dot(N, d, w), del = ora_dot(N, d, rt);
  const double den = rho1 > 0.0 ? bet - gam * gam / rho1 : 0.0;
  const double a2 = den > 0.0 ? del / den : 0.0;
  const double a = a1 - (rho1 > 0.0 ? gam * a2 / rho1 : 0.0);
  for (size_t P = 0; P < N; P++) out[P] = a * c[P] + a2 * d[P];
  free(c); free(v); free(d); free(w); free(rt);
}

void ora_ac_precond(ora_ac *ac, int l, const double *r, double *z, int nu1,
                    int nu2, double omega, double cfac) {
  ora_ac_precond_s(ac, l, r, z, nu1, nu2, omega, cfac, 0);
}

void ora_ac_precond_s(ora_ac *ac, int l, const double *r, double *z, int nu1,
                      int nu2, double omega, double cfac, int smoother) {
  const ora_mat *A = &ac->A[l];
  const size_t N = mat_n(A);
  if (l == ac->nlev - 1) {
    const double *G = ac->G;
    for (size_t i = 0; i < N; i++) {
      double s = 0.0;
      for (size_t j = 0; j < N; j++) s += G[i * N + j] * r[j];
      z[i] = s;
    }
    return;
  }
  double *t = ac->res[l];
  double *ld = NULL, *ly = NULL;
  if (smoother == 1) {
    ld = dalloc(N);
    ly = dalloc((size_t)A->nz);
  }
  /* first sweep from zero: e = omega r / c (point) or omega T^-1 r (z-line) */
  if (smoother == 1 && nu1 > 0) {
    zline_solve(A, r, ld, ly);
    for (size_t P = 0; P < N; P++) z[P] = omega * ld[P];
  } else {
    for (size_t P = 0; P < N; P++) z[P] = nu1 > 0 ? omega * r[P] / A->c[P] : 0.0;
  }
  for (int s = 1; s < nu1; s++) {
    if (smoother == 1) zline_sweep(A, r, z, t, ld, ly, omega);
    else jacobi_sweep(A, r, z, t, omega);
  }
  /* residual and restriction (block sum, inactive 0).  Unlike the
   * reference's stand-alone cycle (additive_restriction.cpp:151-174) the
   * mean is not removed here: the residual sums to zero up to rounding
   * (constants are in the left null space), the Jacobi smoothers do not
   * drift on it, and the coarsest operator G projects it out. */
  ora_resid(A, r, z, t);
  const int nX = ac->nx[l + 1], nY = ac->ny[l + 1], nZ = ac->nz[l + 1];
  const int ny = A->ny, nz = A->nz;
  const int CI = A->nx / nX, CJ = ny / nY, CK = nz / nZ;
  double *rH = ac->f[l + 1], *eH = ac->phi[l + 1];
  const size_t NH = (size_t)nX * nY * nZ;
  const int *fl = ac->aflag[l + 1];
  const int *ff = ac->aflag[l]; /* solid fine cells take no part in R = P^T */
  for (int I = 0; I < nX; I++)
    for (int J = 0; J < nY; J++)
      for (int K = 0; K < nZ; K++) {
        double s = 0.0;
        for (int i = 0; i < CI; i++)
          for (int j = 0; j < CJ; j++)
            for (int k = 0; k < CK; k++) {
              size_t P = IDX(I * CI + i, J * CJ + j, K * CK + k, ny, nz);
              if (ff[P]) s += t[P];
            }
        rH[IDX(I, J, K, nY, nZ)] = s;
      }
  for (size_t Q = 0; Q < NH; Q++)
    if (!fl[Q]) rH[Q] = 0.0;
  double cf = cfac;
  if (l + 1 < ac->nlev - 1 && ((kcycle_mask >> (l + 1)) & 1)) {
    ac_kstep(ac, l + 1, rH, eH, nu1, nu2, omega, cfac, smoother);
    cf = 1.0;  /* the Krylov steps choose the scaling */
  } else {
    ora_ac_precond_s(ac, l + 1, rH, eH, nu1, nu2, omega, cfac, smoother);
  }
  for (int I = 0; I < nX; I++)
    for (int J = 0; J < nY; J++)
      for (int K = 0; K < nZ; K++) {
        double v = cf * eH[IDX(I, J, K, nY, nZ)];
        for (int i = 0; i < CI; i++)
          for (int j = 0; j < CJ; j++)
            for (int k = 0; k < CK; k++) {
              size_t P = IDX(I * CI + i, J * CJ + j, K * CK + k, ny, nz);
              if (ff[P]) z[P] += v;
            }
      }
  for (int s = 0; s < nu2; s++) {
    if (smoother == 1) zline_sweep(A, r, z, t, ld, ly, omega);
    else jacobi_sweep(A, r, z, t, omega);
  }
  free(ld);
  free(ly);
}

static double nrm2(size_t n, const double *x) { return sqrt(ora_dot(n, x, x)); }

/* z -= mean(z) over the active cells: keeps the search directions out of the
 * constant null space of the Neumann operator */
static void project_mean(const ora_ac *ac, double *z) {
  const size_t N = mat_n(&ac->A[0]);
  const int *fl = ac->aflag[0];
  double s = 0.0;
  long cnt = 0;
  for (size_t P = 0; P < N; P++)
    if (fl[P]) { s += z[P]; cnt++; }
  if (!cnt) return;
  s /= (double)cnt;
  for (size_t P = 0; P < N; P++)
    if (fl[P]) z[P] -= s;
}

/* the outer CG's beta: 0 Fletcher-Reeves, rho_new / rho (cg.cpp:93); 1
 * flexible, -z.A p / p.A p (Y. Notay, "Flexible conjugate gradients", SIAM J.
 * Sci. Comput. 22 (2000) 1444-1460), which makes each new direction
 * A-orthogonal to the last one whatever the preconditioner did */
static int flex_beta = 0;
void ora_set_flexible(int f) { flex_beta = f; }

/* test hook: orth[k] = |p_k+1 . A p_k| / sqrt(p_k+1 . A p_k+1  p_k . A p_k)
 * for the directions of the next ora_acpcg_s runs (NULL: off) */
static double *orth_buf = NULL;
static int orth_cap = 0;
void ora_set_orth_record(double *buf, int cap) {
  orth_buf = buf;
  orth_cap = cap;
}

int ora_acpcg(ora_ac *ac, double *x, const double *b, double rtol, int maxit,
              int nu1, int nu2, double omega, double cfac, double *hist,
              int *iters) {
  return ora_acpcg_s(ac, x, b, rtol, maxit, nu1, nu2, omega, cfac, 0, hist, iters);
}

int ora_acpcg_s(ora_ac *ac, double *x, const double *b, double rtol, int maxit,
                int nu1, int nu2, double omega, double cfac, int smoother,
                double *hist, int *iters) {
  const ora_mat *A = &ac->A[0];
  const size_t N = mat_n(A);
  if (!ac->G) ora_ac_coarse_dense(ac);
  double *r = dalloc(N), *z = dalloc(N), *p = dalloc(N), *q = dalloc(N);
  ora_resid(A, b, x, r);
  double r0 = nrm2(N, r);
  int it = 0, conv = 0;
  if (hist) hist[0] = 1.0;
  if (r0 == 0.0) { conv = 1; goto done; }
  ora_ac_precond_s(ac, 0, r, z, nu1, nu2, omega, cfac, smoother);
  project_mean(ac, z);
  memcpy(p, z, N * 8);
  double rho = ora_dot(N, r, z);
  for (it = 1; it <= maxit; it++) {
    ora_spmv(A, p, q);
    const double pq = ora_dot(N, p, q);
    double alpha = rho / pq;
    for (size_t P = 0; P < N; P++) x[P] += alpha * p[P];
    for (size_t P = 0; P < N; P++) r[P] -= alpha * q[P];
    double rel = nrm2(N, r) / r0;
    if (hist) hist[it] = rel;
    if (rel < rtol) { conv = 1; break; }
    ora_ac_precond_s(ac, 0, r, z, nu1, nu2, omega, cfac, smoother);
    project_mean(ac, z);
    double rho_new = ora_dot(N, r, z);
    double beta = flex_beta ? (pq != 0.0 ? -ora_dot(N, z, q) / pq : 0.0) : rho_new / rho;
    rho = rho_new;
    for (size_t P = 0; P < N; P++) p[P] = z[P] + beta * p[P];
    if (orth_buf && it - 1 < orth_cap) {
      double *w = dalloc(N);
      ora_spmv(A, p, w);
      orth_buf[it - 1] = fabs(ora_dot(N, p, q)) / sqrt(ora_dot(N, p, w) * pq);
      free(w);
    }
  }
  if (it > maxit) it = maxit;
done:
  if (iters) *iters = it;
  free(r); free(z); free(p); free(q);
  return conv;
}

/* BiCGS (bicgs.cpp:9-229 steps and stopping rules) with the V-cycle of the
 * hierarchy as its preconditioner in place of the diagonal: not a reference
 * configuration (its BiCGS takes a diagonal or IC preconditioner) */
int ora_acbicgs(ora_ac *ac, double *x, const double *b, const double *dV, int mini,
                int maxit, double res_rat, double res_tol, double scale, int stalecount,
                int nu1, int nu2, double omega, double cfac, double *hist,
                ora_result *out) {
  if (!ac->G && ora_ac_coarse_dense(ac) < 0) return -1; /* singular coarsest level */
  const bicgs_prec M = {NULL, ac, nu1, nu2, omega, cfac};
  return bicgs_run(&ac->A[0], &M, x, b, dV, mini, maxit, res_rat, res_tol, scale,
                   stalecount, hist, out);
}

int ora_pcg_rel(const ora_mat *A, double *x, const double *b, double rtol,
                int maxit, double *hist, int *iters) {
  const size_t N = mat_n(A);
  double *r = dalloc(N), *z = dalloc(N), *p = dalloc(N), *q = dalloc(N);
  ora_resid(A, b, x, r);
  double r0 = nrm2(N, r);
  int it = 0, conv = 0;
  if (hist) hist[0] = 1.0;
  if (r0 == 0.0) { conv = 1; goto done; }
  for (size_t P = 0; P < N; P++) z[P] = r[P] / A->c[P];
  memcpy(p, z, N * 8);
  double rho = ora_dot(N, r, z);
  for (it = 1; it <= maxit; it++) {
    ora_spmv(A, p, q);
    double alpha = rho / ora_dot(N, p, q);
    for (size_t P = 0; P < N; P++) x[P] += alpha * p[P];
    for (size_t P = 0; P < N; P++) r[P] -= alpha * q[P];
    double rel = nrm2(N, r) / r0;
    if (hist) hist[it] = rel;
    if (rel < rtol) { conv = 1; break; }
    for (size_t P = 0; P < N; P++) z[P] = r[P] / A->c[P];
    double rho_new = ora_dot(N, r, z);
    double beta = rho_new / rho;
    rho = rho_new;
    for (size_t P = 0; P < N; P++) p[P] = z[P] + beta * p[P];
  }
  if (it > maxit) it = maxit;
done:
  if (iters) *iters = it;
  free(r); free(z); free(p); free(q);
  return conv;
}
