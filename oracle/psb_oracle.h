/*
 * psb_oracle.h -- CPU restatement (FP64, plain C) of PSI-BOIL's pressure-Poisson
 * solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by the
 * tests, by the smoke check and by the CPU-baseline leg of the benchmark, and
 * never by the product library.  Nothing in psi-boil_amd/ links or calls it.
 *
 * Layout of every field: interior cells only, index (i*ny + j)*nz + k, k
 * fastest -- the reference's [i][j][k] order without its BW=3 ghost padding
 * (Src/Global/global_malloc.cpp:35-56, Src/Field/Scalar/scalar.cpp:77-93).
 * Neighbours outside the interior are read as 0.0; the reference reads ghost
 * values there, but every coefficient on a Neumann wall is 0, so the product
 * is the same (Src/Equation/Centered/centered_system_bnd.cpp:19-61).
 *
 * Each function cites the reference file:line it follows.  Which checks pin
 * which function is listed at the top of psb_oracle.c.
 */
#ifndef PSB_ORACLE_H
#define PSB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 7-diagonal matrix: c, w, e, s, n, b, t (Src/Matrix/matrix.h:23-54).
 * Off-diagonals positive; (Ax)_P = c_P x_P - sum a_nb x_nb. */
typedef struct {
  int nx, ny, nz;
  double *c, *w, *e, *s, *n, *b, *t;
} ora_mat;

typedef struct {
  int iters;       /* reference's printed count: i+1 (cg.cpp:178)           */
  int converged;   /* return value of solve()                               */
  int reason;      /* 0 none/maxiter, 1 res_tol, 2 res_rat, 3 stale,
                      4 breakdown rho, 5 small s, 6 small omega, 7 res0<tol  */
  double res0, res;
} ora_result;

/* ---- splitmix64 and the synthetic case generator ----------------------- */
uint64_t ora_splitmix64(uint64_t *state);
double   ora_u01(uint64_t *state);

/* Spheres: centres uniform in [r, L-r] per axis, radius uniform in [rmin,rmax].
 * out: 4*nsph doubles (cx, cy, cz, r).  Draw order: r, cx, cy, cz.          */
void ora_gen_spheres(uint64_t seed, int nsph, double rmin, double rmax,
                     double Lx, double Ly, double Lz, double *out);

/* Cell-centre metrics from interior nodes xn[0..n] with the reference's
 * default BndGrid::wall() ghost rule (Src/Grid/grid1d_correct_boundaries.cpp).
 * xc[n], dxc[n], dxw[n] (=dxn(i)), dxe[n] (=dxn(i+1)).                      */
void ora_metrics(int n, const double *xn, double *xc, double *dxc,
                 double *dxw, double *dxe);

/* Cell density from spheres: inside <=> |x_c - c| < r (any sphere).  rho is
 * (nx+2)(ny+2)(nz+2) with a one-cell ghost layer holding the adjacent
 * interior value (Neumann colour function).                                  */
void ora_density(int nx, int ny, int nz, const double *xc, const double *yc,
                 const double *zc, int nsph, const double *sph,
                 double rho_in, double rho_out, double *rho);

/* Pressure::discretize with all six walls Neumann and an optional voxel
 * "off" mask (1 = cell inside a solid).  Src/Equation/Centered/Pressure/
 * pressure_discretize.cpp:11-151, Src/Matter/property_value.cpp:18-32,
 * Src/Equation/Centered/centered_system_bnd.cpp:19-61.  ci may be NULL.      */
void ora_discretize(int nx, int ny, int nz,
                    const double *dxc, const double *dxw, const double *dxe,
                    const double *dyc, const double *dys, const double *dyn,
                    const double *dzc, const double *dzb, const double *dzt,
                    const double *rho, const uint8_t *off,
                    ora_mat *A, double *ci, double *dV);
/* ora_discretize with the pressure's boundary conditions per side (imin,
 * imax, jmin, jmax, kmin, kmax; centered_system_bnd.cpp): btype 0 the
 * Neumann group (wall, symmetry, outlet, neumann), 1 the Dirichlet group
 * (dirichlet, inlet); bval the side's value; fbnd (compact) receives the
 * boundary source.  btype NULL: walls everywhere (ora_discretize).         */
void ora_discretize_bnd(int nx, int ny, int nz, const double *dxc, const double *dxw,
                        const double *dxe, const double *dyc, const double *dys,
                        const double *dyn, const double *dzc, const double *dzb,
                        const double *dzt, const double *rho, const uint8_t *off,
                        const int *btype, const double *bval, ora_mat *A, double *ci,
                        double *dV, double *fbnd);

/* Pressure::update_rhs with dti = 1: b = (u_w-u_e)dSx + (v_s-v_n)dSy +
 * (w_b-w_t)dSz (pressure_update_rhs.cpp:27-52), b = 0 on off cells, then the
 * zero-mean shift over active cells that the benchmark cases add.
 * u: (nx+1)*ny*nz, v: nx*(ny+1)*nz, w: nx*ny*(nz+1), face index along its
 * own axis outermost-first like the cells.                                  */
void ora_rhs(int nx, int ny, int nz, const double *dxc, const double *dyc,
             const double *dzc, const double *u, const double *v,
             const double *w, const uint8_t *off, int zero_mean, double *b);
/* ora_rhs plus fbnd (NULL: none) before the optional mean removal, as
 * update_rhs adds fext + fbnd (pressure_update_rhs.cpp:135)               */
void ora_rhs_bnd(int nx, int ny, int nz, const double *dxc, const double *dyc,
                 const double *dzc, const double *u, const double *v, const double *w,
                 const uint8_t *off, const double *fbnd, int zero_mean, double *b);

/* Face velocities uniform(-1,1), wall-normal boundary faces 0.
 * Draw order: all u faces, then v, then w, each in storage order.           */
void ora_face_velocity(uint64_t seed, int nx, int ny, int nz,
                       double *u, double *v, double *w);

/* Pressure::ghost (pressure_ghost.cpp:10-190): fext += the ghost-fluid
 * surface-tension source of every face the colour function c crosses at 0.5;
 * c, kappa, rho with one ghost layer (as ora_discretize's rho), fext
 * compact; solid cells (off) skipped.                                       */
void ora_ghost(int nx, int ny, int nz, const double *dxc, const double *dxw,
               const double *dxe, const double *dyc, const double *dys,
               const double *dzc, const double *dzb, const double *dzt,
               const double *c, const double *kappa, const double *rho, double sigma,
               const uint8_t *off, double *fext);

/* Momentum::project (momentum_project.cpp:10-76) on the interior faces of
 * u, v, w (layout as ora_rhs) with the pressure p and the ghosted density
 * of ora_discretize; off != NULL selects the body form (faces between two
 * fluid cells only).  mask: bit 0 u, bit 1 v, bit 2 w.                      */
void ora_project(int nx, int ny, int nz, const double *dxw, const double *dys,
                 const double *dzb, const double *rho, const uint8_t *off,
                 const double *p, double dt, int mask, double *u, double *v,
                 double *w);

/* Summation order of ora_dot / ora_dot_voldiv_avg: 0 (default) sequential
 * like the reference, b > 0 blocks of b terms then the block sums (test
 * infrastructure: sensitivity of a result to the order of its sums).        */
void ora_set_sum_block(int b);

/* ---- vector kernels ----------------------------------------------------- */
void   ora_spmv(const ora_mat *A, const double *x, double *y);
void   ora_resid(const ora_mat *A, const double *b, const double *x, double *r);
double ora_dot(size_t n, const double *x, const double *y);
double ora_dot_voldiv_avg(size_t n, const double *x, const double *y,
                          const double *dV);

/* ---- Krylov / iterative solvers (reference semantics) ------------------- */
int ora_cg(const ora_mat *A, double *x, const double *b, const double *dV,
           int mini, int maxit, double res_rat, double res_tol, double scale,
           int stalecount, double *hist, ora_result *out);
/* CG with a preconditioner: ORA_PREC_DI / _IC0 / _IC2 as Prec::di / ic0 /
 * ic2 (preconditioner.h:16-20), ORA_PREC_RBSGS the red-black symmetric
 * Gauss-Seidel the library runs in their place; ora_prec_apply: z = M^-1 r */
enum { ORA_PREC_DI = 0, ORA_PREC_IC0 = 1, ORA_PREC_IC2 = 2, ORA_PREC_RBSGS = 10 };
int ora_cg_prec(const ora_mat *A, int prec, double *x, const double *b, const double *dV,
                int mini, int maxit, double res_rat, double res_tol, double scale,
                int stalecount, double *hist, ora_result *out);
void ora_prec_apply(const ora_mat *A, int kind, const double *r, double *z);
int ora_bicgs(const ora_mat *A, double *x, const double *b, const double *dV,
              int mini, int maxit, double res_rat, double res_tol, double scale,
              int stalecount, double *hist, ora_result *out);
int ora_jacobi(const ora_mat *A, double *x, const double *b, int mini,
               int maxit, double res_rat, double res_tol, ora_result *out);

/* ---- additive-correction multigrid -------------------------------------- */
int  ora_smallest_factor(int n);                 /* Src/Global/global_factor.cpp */
/* Level sizes by Src/Domain/domain_coarsen.cpp:6-112 (rule 0), by the
 * benchmark rule "halve a direction while it stays >= 8" (rule 1) or by
 * "halve while it stays >= 4" (rule 2, the batched solver's hierarchy).
 * Returns the number of levels; nl{x,y,z}[0] = n{x,y,z}.                    */
int  ora_levels(int rule, int nx, int ny, int nz, int maxlev,
                int *lx, int *ly, int *lz);

typedef struct {
  int nlev;
  int nx[24], ny[24], nz[24];
  ora_mat A[24];
  int *aflag[24];
  double *dV[24];
  double *phi[24], *f[24], *res[24];
  /* dense coarsest-level operator for the preconditioned form (NULL until
   * ora_ac_coarse_dense is called) */
  double *G;
} ora_ac;

/* Allocate levels, copy the fine matrix, coarsen aflag and the Galerkin
 * system for every level (additive.cpp:8-48, additive_coarsen_flag.cpp,
 * additive_coarsen_system.cpp).  aflag0 may be NULL (all active).           */
int  ora_ac_build(ora_ac *ac, int rule, const ora_mat *A0, const double *dV0,
                  const uint8_t *off0, int maxlev);
void ora_ac_free(ora_ac *ac);
void ora_coarsen_flag(const ora_ac *ac, int l);       /* level l-1 -> l       */
void ora_coarsen_system(const ora_ac *ac, int l);     /* level l-1 -> l       */
void ora_restriction(const ora_ac *ac, int l);        /* level l   -> l+1     */
void ora_interpolation(const ora_ac *ac, int l);      /* level l   -> l-1     */

/* AC::cycle(none, V, toler, factor, mv, ms) with CG-di smoothers on every
 * level (additive_cycle.cpp, additive_cycle_functions.cpp,
 * additive_recursive.cpp:4-34).  phi0/f0 are level 0's phi and fnew.
 * hist: residual after each cycle (max_cyc+1 entries, [0] = initial).       */
int  ora_ac_vcycle(ora_ac *ac, double *phi0, const double *f0, double toler,
                   double factor, int max_cyc, int min_cyc, const int mv[3],
                   double dt, double *hist, int *ncyc);

/* AC::cycle (additive_cycle.cpp:4-99): the loop's cycle shape, an optional
 * full-cycle initialisation (additive_recursive.cpp:127-163) and the
 * controls of additive.h:97-142.  Cycle::flex (loop only) with a bound of
 * max_visits level visits per cycle (the reference has none: its recursion
 * ends only when a level's smoother meets its ResRat); a cycle that reaches
 * it ends the run, ora_ac_cycle returns -1 and phi holds the previous
 * cycle's result. */
enum { ORA_CYC_NONE = -1, ORA_CYC_V = 0, ORA_CYC_F = 1, ORA_CYC_W = 2, ORA_CYC_FLEX = 3 };
typedef struct {
  double toler, factor;          /* ResTol, ResRat of the loop               */
  int max_cyc, min_cyc;
  int mv[3];                     /* MaxIter pre / re- / post-smoothing       */
  int ms[3];                     /* stale counts of those solves (-1: off)   */
  double dt;
  int init, loop;                /* ORA_CYC_*                                */
  int stop_if_div, priority_min_cyc, use_linf;
  double resrat, restol;         /* level solves (AC::resrat / restol)       */
  int max_visits;                /* Cycle::flex: level visits per cycle      */
  int solver;                    /* level solves: 0 CG, 1 BiCGS (diagonal
                                    preconditioner both, the equation's solver
                                    of additive_cycle_functions.cpp:89-107) */
} ora_cycle;
void ora_cycle_defaults(ora_cycle *P);
/* hist, hist_linf: max_cyc + 2 entries; *nhist receives the count.          */
int  ora_ac_cycle(ora_ac *ac, double *phi0, const double *f0, const ora_cycle *P,
                  double *hist, double *hist_linf, int *nhist, int *ncyc);

/* ---- the benchmark's composition (not in the reference) ----------------- */
/* Dense operator of the coarsest level: G = Pi * (A_red)^-1 * Pi over active
 * cells, Pi the mean projector, A_red the matrix with its last pivot pinned
 * (A^-1 itself when an active row of the finest matrix does not sum to
 * zero: a non-singular operator, Dirichlet sides);
 * for a non-symmetric level (7 arrays) G = A^-1 on the active cells by LU
 * with partial pivoting.  Returns the level's cell count, -1 for a singular
 * non-symmetric level (G left unset). */
int  ora_ac_coarse_dense(ora_ac *ac);
/* One V-cycle as a preconditioner z = M^-1 r: nu1 damped-Jacobi sweeps from
 * zero, residual restriction (block sums), recursion, piecewise-constant
 * prolongation scaled by cfac, nu2 sweeps; coarsest level z = G r.          */
void ora_ac_precond(ora_ac *ac, int l, const double *r, double *z, int nu1,
                    int nu2, double omega, double cfac);
/* Same with a choice of smoother: 0 point Jacobi (as above), 1 z-line
 * Jacobi, e += omega T^-1 (r - A e) with T the tridiagonal part of A along
 * k of each (i, j) column (Thomas algorithm).                              */
void ora_ac_precond_s(ora_ac *ac, int l, const double *r, double *z, int nu1,
                      int nu2, double omega, double cfac, int smoother);
/* K-cycle: levels (bit l = level l, 0 < l < coarsest) whose coarse
 * correction in ora_ac_precond* is two flexible-CG steps preconditioned by
 * the cycle below instead of cfac times one cycle; 0 (default): V-cycle.    */
void ora_set_kcycle(int mask);
/* outer CG beta of ora_acpcg_s: 0 Fletcher-Reeves (default), 1 flexible */
void ora_set_flexible(int f);
/* test hook: A-orthogonality of successive directions of ora_acpcg_s */
void ora_set_orth_record(double *buf, int cap);
/* x = T^-1 t column by column (the z-line smoother's solve).               */
void ora_zline_solve(const ora_mat *A, const double *t, double *x);
/* PCG with the V-cycle preconditioner, its output projected to zero mean over
 * the active cells; stops on ||r||2/||r0||2 < rtol.
 * hist[i] = relative residual after iteration i (hist[0] = 1).             */
int  ora_acpcg(ora_ac *ac, double *x, const double *b, double rtol, int maxit,
               int nu1, int nu2, double omega, double cfac, double *hist,
               int *iters);
int  ora_acpcg_s(ora_ac *ac, double *x, const double *b, double rtol, int maxit,
                 int nu1, int nu2, double omega, double cfac, int smoother,
                 double *hist, int *iters);
/* BiCGS with the reference's steps and stopping rules (bicgs.cpp:9-229, as
 * ora_bicgs) and the V-cycle (nu1, nu2, omega, cfac) as its preconditioner */
int  ora_acbicgs(ora_ac *ac, double *x, const double *b, const double *dV, int mini,
                 int maxit, double res_rat, double res_tol, double scale, int stalecount,
                 int nu1, int nu2, double omega, double cfac, double *hist,
                 ora_result *out);
/* Diagonal PCG with the same relative-2-norm stopping rule. */
int  ora_pcg_rel(const ora_mat *A, double *x, const double *b, double rtol,
                 int maxit, double *hist, int *iters);

#ifdef __cplusplus
}
#endif
#endif
