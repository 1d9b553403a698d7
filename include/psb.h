/*
 * psb.h -- C ABI of the MI355X pressure-Poisson solver (libpsb.so).
 *
 * Drop-in for the solve path of PSI-BOIL: the 7-diagonal Matrix that
 * Pressure::discretize builds (Src/Matrix/matrix.h:23-54), solved by
 *   Linear::solve(A, x, b, MinIter, MaxIter, name, ResRat, ResTol, scale,
 *                 stalecount, precform)        Src/Solver/Linear/linear.h:46-56
 * with CG or BiCGS and diagonal preconditioning (Src/Solver/Linear/Krylov/
 * cg.cpp:10-182, bicgs.cpp:9-229, Src/Solver/Preconditioner/
 * preconditioner_form.cpp:4-13), or with the additive-correction multigrid
 * (Src/Solver/Additive/) as a CG preconditioner.
 *
 * Arrays: plain pointers, cell order [i][j][k] with k fastest, like the
 * reference's Scalar storage (Src/Field/Scalar/scalar.cpp:77-93).  Host
 * arrays may carry the reference's ghost padding (psb_layout.ghost = 3, its
 * BW, Src/Global/global_buffer_width.cpp:4); device arrays are always the
 * compact interior.  Every entry point returns a psb_status; none throws.
 * A context is not thread-safe: calls on one context are serialised by the
 * caller.  Every call returns when its device work is complete (host and
 * device buffers alike); a context's stream is a blocking one, so device
 * arrays the caller wrote on the null stream are complete before it reads
 * them (work on the caller's own non-blocking streams must be synchronized
 * by the caller).
 */
#ifndef PSB_H
#define PSB_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct psb_ctx psb_ctx;

/* Status codes.  The reference reports failure only as `false` from
 * solve() plus a log line; the negative codes have no counterpart there. */
enum psb_status {
  PSB_OK = 0,          /* converged (solve() returned true)                  */
  PSB_MAXITER = 1,     /* MaxIter reached (solve() returned false)           */
  PSB_STALE = 2,       /* stalecount window without improvement (cg.cpp:138) */
  PSB_BREAKDOWN = 3,   /* BiCGS rho/omega/s breakdown (bicgs.cpp:89,134,188) */
  PSB_DIVERGED = 4,    /* refinement or cycle stopped reducing the residual  */
  PSB_NAN = 5,         /* non-finite residual                                */
  PSB_EINVAL = -1,     /* bad argument, shape or unsupported matrix           */
  PSB_EHIP = -2,       /* HIP runtime error                                  */
  PSB_ENOMEM = -3,     /* device allocation failed                           */
  PSB_ENOTREADY = -4   /* psb_set_matrix not called                          */
};

enum psb_method { PSB_CG = 0, PSB_BICGS = 1 };
/* PSB_PREC_DIAG: Prec::di (preconditioner_form.cpp:4-13).  PSB_PREC_AC: the
 * additive-correction V-cycle, for CG (any precision, symmetric matrix) or
 * for BiCGS (FP64, or PSB_MIXED for an FP32 cycle on a symmetric matrix with
 * the BiCGS itself in FP64; the V-cycle on every level whatever ac_kcycle
 * says, as BiCGS needs the same operator at each application; a non-symmetric matrix
 * keeps 7 arrays on every level, its coarsest level inverted by LU, and a
 * singular one -- no time term -- is PSB_EINVAL).  PSB_PREC_SGS: red-black symmetric
 * Gauss-Seidel, M = (D - L) D^-1 (D - U) with the even cells ((i+j+k) % 2 = 0)
 * ordered first, the parallel replacement for Prec::ic0 / ic2 / ic3 (whose
 * triangular sweeps, preconditioner_solve.cpp:14-110, run cell after cell);
 * CG and FP64 only.  Not the same preconditioner: on the fixture cases CG
 * needs about 1.5-1.8 times the iterations of Prec::ic2 and half those of
 * Prec::di, each iteration in parallel. */
enum psb_precond { PSB_PREC_DIAG = 0, PSB_PREC_AC = 1, PSB_PREC_SGS = 2 };
enum psb_precision { PSB_FP64 = 0, PSB_MIXED = 1 };
enum psb_norm {
  PSB_NORM_REF = 0,   /* sqrt(sum (r/dV)^2)/N/scale, cg.cpp:49,121 with
                         Scalar::dot_voldiv_avg (scalar.h:317-326)          */
  PSB_NORM_L2REL = 1  /* ||r||_2 compared with res_rat * ||r_0||_2          */
};

typedef struct {
  int nx, ny, nz;   /* interior cells                                        */
  int ghost;        /* ghost cells on each side of host arrays (0 or 3)      */
} psb_layout;

/* Host arrays of the seven diagonals (positive off-diagonals; (Ax)_P =
 * c x_P - sum a_nb x_nb).  Couplings across the domain boundary must be 0,
 * as Centered::create_system_bnd leaves them (centered_system_bnd.cpp:19-61).
 * If e/n/t mirror w/s/b of the neighbour bit for bit (the pressure matrix,
 * pressure_discretize.cpp:24-57), only c, w, s, b are stored on the device. */
typedef struct {
  const double *c, *w, *e, *s, *n, *b, *t;
} psb_matrix7;

/* Stopping criteria and method selection: MinIter, MaxIter, ResRat, ResTol,
 * scale, stalecount, precform of Linear::solve, plus the choices the
 * reference makes at construction time (CG/BiCGS class, Prec::di()).
 * Fill it with psb_stop_defaults, which sets struct_size: every entry point
 * that takes a psb_stop returns PSB_EINVAL when struct_size is not the
 * sizeof(psb_stop) the library was built with (a caller compiled against
 * another version of this header). */
typedef struct {
  int struct_size;    /* sizeof(psb_stop), set by psb_stop_defaults          */
  int min_iter;       /* MinIter  (Src/Ravioli/maxiter.h:31, default 1)      */
  int max_iter;       /* MaxIter  (maxiter.h:17, default 20)                 */
  double res_rat;     /* ResRat   (resrat.h:26, default 0.01)                */
  double res_tol;     /* ResTol   (restol.h:26, default 1e-15)               */
  double scale;       /* scale    (linear.h:53, default 1.0)                 */
  int stalecount;     /* stalecount (linear.h:54, default -1 = off; at most
                         64, larger values are PSB_EINVAL)                   */
  int precform;       /* precform (linear.h:55); 1/c is formed on the fly    */
  int method;         /* psb_method                                          */
  int precond;        /* psb_precond                                         */
  int precision;      /* psb_precision                                       */
  int norm;           /* psb_norm                                            */
  /* additive-correction preconditioner (PSB_PREC_AC)                       */
  int ac_pre, ac_post;   /* damped Jacobi sweeps before / after (2, 2)      */
  double ac_omega;       /* Jacobi damping (0.9)                            */
  double ac_cfac;        /* coarse-correction factor (2.0)                  */
  /* iterative refinement (PSB_MIXED)                                       */
  double inner_rtol;     /* inner FP32 solve: relative L2 reduction (1e-4);
                            with inner_adapt = 2 the drop between updates   */
  int max_outer;         /* refinement steps (30)                            */
  int graph_chunk;       /* iterations per captured graph launch (8)         */
  int inner_adapt;       /* 0: restarted refinement, fixed inner_rtol (default);
                            1: restarted, each inner solve aims at the outer
                            target, max(inner_rtol, 0.5 target / current);
                            2: one FP32 CG, never restarted, with reliable
                            updates: whenever its recursive residual falls
                            by inner_rtol below the last true one, x += d in
                            FP64 and r = b - A x (FP64) replaces it; the stop
                            tests use that true residual in the chosen norm */
  int ac_coarsen;        /* AC levels: 0 halve every direction that stays
                            >= 8 (default); 1 halve x and y first, z only
                            once x and y stop (grids stretched in z)       */
  int ac_store;          /* PSB_MIXED level matrices: 0 FP32 (default);
                            1 FP16 copy scaled by a power of two per level
                            for the smoother (half its matrix bytes; the CG
                            operator and the residual updates keep FP32 /
                            FP64); 2 the same copy also for the FP32 CG
                            operator on level 0 (its q = A p reads 8 instead
                            of 16 B per cell; the FP64 reliable updates
                            still measure the true residual)                */
  int ac_fuse;           /* PSB_MIXED: 1 runs the two post-sweeps of a level
                            in one pass (2.5-D marching tile) where the
                            level's shape allows; 0 one kernel per sweep
                            (default: faster on MI355X today; same z)      */
  int ac_kcycle;         /* K-cycle levels, bit l for level l (level 1 is the
                            first coarse level; bit 0 must be 0): there the
                            coarse correction is two flexible-CG steps
                            preconditioned by the cycle below (Notay and
                            Vassilevski 2008) instead of cfac times one
                            cycle.  Bits of the coarsest level or of levels
                            the hierarchy does not have are ignored.
                            -1: chosen from the hierarchy: halving every
                            direction (ac_coarsen = 0), every level between
                            the finest and the coarsest with 16^3 to 64^3
                            cells; halving x and y first (ac_coarsen = 1),
                            levels 1-4 (the benchmark's setting).
                            0 (default): the V-cycle.                       */
  int ac_flexible;       /* outer CG's beta: 1 flexible (Polak-Ribiere form,
                            beta = -z_new.A p / p.A p, which keeps successive
                            directions A-orthogonal when the preconditioner
                            varies from one application to the next, as the
                            K-cycle does); 0 Fletcher-Reeves, rho_new / rho
                            (cg.cpp:93); -1 (default): flexible exactly when
                            the K-cycle is on                               */
} psb_stop;

typedef struct {
  int iterations;    /* reference count i+1 (cg.cpp:178); inner total (IR)   */
  int outer;         /* refinement steps / FP64 updates (PSB_MIXED), else 0 */
  int converged;     /* solve()'s return value                              */
  int status;        /* psb_status                                          */
  double res0, res;  /* initial / final residual in the chosen norm         */
  double rel_l2;     /* true ||b - A x||_2 / ||b||_2 at exit (FP64)          */
} psb_result;

/* Fill *s with the reference's defaults (CG, diagonal, FP64, reference
 * norm, MinIter 1, MaxIter 20, ResRat 0.01, ResTol 1e-15, scale 1). */
void psb_stop_defaults(psb_stop *s);
const char *psb_status_string(int status);

/* Linear(const Domain&, const Prec&) / CG(...) / BiCGS(...) construction:
 * Src/Solver/Linear/linear.cpp:13-35, Krylov/cg.h:17-18.  Allocates device
 * buffers for the grid and a HIP stream.                                     */
int psb_create(psb_ctx **ctx, const psb_layout *layout);
int psb_destroy(psb_ctx *ctx);

/* Cell sizes per direction (Domain::dxc/dyc/dzc, domain.h:81-88), used for
 * dV = dxc*dyc*dzc in the reference norm (domain_surfaces_volumes.cpp:62-63).
 * NULL pointers mean unit cells.                                            */
int psb_set_geometry(psb_ctx *ctx, const double *dxc, const double *dyc,
                     const double *dzc);

/* Upload A (host arrays in the context's layout), build the FP32 copy and
 * the AC hierarchy when needed.  Replaces holding a Matrix& across calls.    */
int psb_set_matrix(psb_ctx *ctx, const psb_matrix7 *A);

/* Off cells (1 = solid, Body::off_p) for the AC hierarchy's active flag
 * (Centered::set_active_flag, additive_coarsen_flag.cpp:4-41).  Optional;
 * takes effect at the next psb_set_matrix.                                  */
int psb_set_active(psb_ctx *ctx, const unsigned char *off);

/* ------------------------------------------------------------------------
 * Matrix and right-hand side built on the device (the solve's producers):
 * Pressure::discretize (pressure_discretize.cpp:11-151, with the wall
 * correction of centered_system_bnd.cpp:19-61) and Pressure::update_rhs
 * (pressure_update_rhs.cpp:16-139, dt = 1).  Only the density (one array)
 * crosses PCIe instead of seven coefficient arrays; the matrix is the one
 * psb_set_matrix would receive from the reference, bit for bit.
 * ------------------------------------------------------------------------ */
typedef struct {   /* per direction, as Domain::dxc/dxw/dxe (domain.h:82-90) */
  const double *dxc, *dxw, *dxe;  /* cell size, distance to the west / east centre */
  const double *dyc, *dys, *dyn;
  const double *dzc, *dzb, *dzt;
} psb_metrics;

/* The pressure's boundary condition on each side of the box (imin, imax,
 * jmin, jmax, kmin, kmax) for psb_assemble_pressure*, as
 * Centered::create_system_bnd applies it (centered_system_bnd.cpp:5-61,
 * 119-160): PSB_BND_NEUMANN for the reference's wall, symmetry, outlet and
 * neumann types (the side's coupling moves out of c; value * face area joins
 * the boundary source), PSB_BND_DIRICHLET for dirichlet and inlet (c keeps
 * the coupling, value * coupling joins the source; the matrix is then
 * non-singular).  The source is added to b by psb_update_rhs_device, as
 * update_rhs adds fbnd (pressure_update_rhs.cpp:135).  Set before the
 * assembly; the default is PSB_BND_NEUMANN with value 0 on every side.     */
enum psb_bnd_type { PSB_BND_NEUMANN = 0, PSB_BND_DIRICHLET = 1 };
typedef struct {
  int type[6];      /* psb_bnd_type per side                                */
  double value[6];  /* the condition's value (the reference's BndCnd value) */
} psb_bnd;
int psb_set_boundary(psb_ctx *ctx, const psb_bnd *bnd);

/* rho: density with one ghost layer, (nx+2)(ny+2)(nz+2) values, [i][j][k]
 * with k fastest (the reference's colour field with Neumann ghosts); off:
 * solid cells (compact interior order, 1 = solid) or NULL.  Replaces
 * psb_set_matrix, psb_set_active and psb_set_geometry (dxc, dyc, dzc).      */
int psb_assemble_pressure(psb_ctx *ctx, const double *rho, const psb_metrics *m,
                          const unsigned char *off);
int psb_assemble_pressure_device(psb_ctx *ctx, const double *d_rho,
                                 const psb_metrics *m, const unsigned char *off);

/* After psb_assemble_pressure*: d_b = (u_w - u_e) dSx + (v_s - v_n) dSy +
 * (w_b - w_t) dSz from the device face velocities u ((nx+1) ny nz), v
 * (nx (ny+1) nz), w (nx ny (nz+1)), 0 on solid cells, plus the boundary
 * source of psb_set_boundary's conditions; zero_mean: minus the mean over
 * the fluid cells (the benchmark's compatible right-hand side).             */
int psb_update_rhs_device(psb_ctx *ctx, const double *d_u, const double *d_v,
                          const double *d_w, int zero_mean, double *d_b);

/* Pressure::ghost (Src/Equation/Centered/Pressure/pressure_ghost.cpp:10-190),
 * the ghost-fluid method's surface-tension source, after
 * psb_assemble_pressure*: d_fext (compact, the caller's Pressure::fext that
 * update_rhs adds to b, pressure_update_rhs.cpp:135) += iphase dS sigma
 * kappa_f / (rho_f d) for every face of a cell that the colour function
 * crosses at 0.5 (iphase -1 on the cell's side below it; rho_f the mean of
 * the two cells' densities, kappa_f their curvatures' harmonic mean when of
 * one sign, else the arithmetic mean; d the centre distance, dys also on the
 * north face as the reference has it).  d_c, d_kappa, d_rho: colour,
 * curvature and density with one ghost layer (the layout of
 * psb_assemble_pressure's rho); solid cells are left unchanged.
 * Bit-identical with the reference.                                         */
int psb_pressure_ghost_device(psb_ctx *ctx, const double *d_c, const double *d_kappa,
                              const double *d_rho, double sigma, double *d_fext);

/* Momentum::project (Src/Equation/Staggered/Momentum/momentum_project.cpp:
 * 10-76), the step after the pressure solve, on device arrays after
 * psb_assemble_pressure*: the interior faces of u, v, w (layout as in
 * psb_update_rhs_device; wall faces unchanged) take the correction
 * (p_nb - p_P) dt / (rho_face d) with d = dxw / dys / dzb.  d_rho: the
 * density the matrix was assembled from (ghosted layout); d_p: the solution
 * of psb_solve_device, used in place.  Without solid cells the rounding of
 * the reference's plain form, with them its body form (faces between two
 * fluid cells only).  mask: bit 0 u, bit 1 v, bit 2 w (the reference's
 * ifull / jfull / kfull); arrays of unselected components may be NULL.
 * The ghost-face update (bnd_update_nooutlet, exchange_all) is the
 * caller's.                                                                 */
int psb_project_device(psb_ctx *ctx, const double *d_rho, const double *d_p,
                       double dt, int mask, double *d_u, double *d_v,
                       double *d_w);

/* ------------------------------------------------------------------------
 * Checkpoints of a cell field in the reference's .bck format
 * (Scalar::save / Scalar::load, Src/Field/Scalar/scalar_save.cpp:20-32,
 * scalar_load.cpp:29-39): the padded extents as three ints, then the padded
 * array (bw ghost layers per side; the reference's BW is 3) as doubles,
 * [i][j][k] with k fastest.  A warm start across time steps or runs reads
 * x back from the file the reference (or psb_save_bck) wrote.
 * ------------------------------------------------------------------------ */
/* name_file (Src/Global/global_name_file.cpp:4-32): name_p<rank>_<step>.bck
 * (rank or step < 0: left out).                                             */
int psb_bck_name(const char *name, int step, int rank, char *out, int cap);
/* x: compact interior nx*ny*nz (host); the ghost layers written hold the
 * nearest interior value.                                                   */
int psb_save_bck(const char *path, int nx, int ny, int nz, int bw,
                 const double *x);
/* Reads a .bck of interior nx*ny*nz with any ghost width (returned in *bw,
 * may be NULL) into the compact x.  PSB_EINVAL on a size mismatch.          */
int psb_load_bck(const char *path, int nx, int ny, int nz, double *x, int *bw);
/* The same for a device array of the context's grid (ghost width 3 on save). */
int psb_save_bck_device(psb_ctx *ctx, const char *path, const double *d_x);
int psb_load_bck_device(psb_ctx *ctx, const char *path, double *d_x);

/* Linear::solve (linear.h:46-56): x (in: initial guess, out: solution) and b
 * are host arrays in the context's layout.                                  */
int psb_solve(psb_ctx *ctx, double *x, const double *b, const psb_stop *stop,
              psb_result *res);

/* Same on device-resident compact arrays (hipMalloc'd by the caller). */
int psb_solve_device(psb_ctx *ctx, double *d_x, const double *d_b,
                     const psb_stop *stop, psb_result *res);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory during a solve (tight serving loops, code that forbids hipMalloc):
 * psb_solve_device for CG with the diagonal preconditioner in FP64 (the only
 * method it accepts; anything else is PSB_EINVAL) that allocates nothing.
 * Its kernels are those of psb_solve_device, launched one by one instead of
 * from a captured graph, and its Krylov vectors, reduction partials and
 * solver state live in a static device array of 64 MiB per device (held by
 * each device the library runs on until the process ends), so the system
 * may have at most 2^21 cells (128^3); larger ones return PSB_EINVAL with a
 * message on stderr.  x (in: initial guess, out: solution) is iterated in
 * place.  The array is shared by every context and call on a device: calls
 * that overlap on one device race.  Results equal psb_solve_device's bit for
 * bit; psb_history reports no history for it.                               */
int psb_solve_device_static(psb_ctx *ctx, double *d_x, const double *d_b,
                            const psb_stop *stop, psb_result *res);

/* Scalar::operator=(A_x) / (y_m_A_x) (scalar_operators.cpp:22-72) on host
 * arrays: y = A x, or y = b - A x when b != NULL.  FP64, bit-exact with the
 * reference. */
int psb_apply(psb_ctx *ctx, const double *x, const double *b, double *y);

/* Residual history of the last solve in its norm: hist[0] = res0, hist[i] =
 * res after iteration i; for PSB_MIXED the FP64 residual before each
 * refinement step (after each update with inner_adapt = 2).  Returns the
 * count.                                                                    */
int psb_history(psb_ctx *ctx, double *hist, int cap);

/* Synthetic cases of the benchmark: variable-density pressure matrix on a
 * uniform (or z-stretched) grid with spheres and an optional solid cylinder,
 * RHS = divergence of a seeded face-velocity field shifted to zero mean.
 * Same arithmetic as Pressure::discretize / update_rhs
 * (pressure_discretize.cpp:11-151, pressure_update_rhs.cpp:16-139).         */
typedef struct {
  int nx, ny, nz;
  int nsph;                       /* random spheres (ignored if centred)     */
  double rmin, rmax;              /* radius range (absolute length)          */
  double rho_in, rho_out;         /* density inside / outside spheres        */
  unsigned long long seed_sph, seed_vel;
  int centred;                    /* 1: one sphere at (cx,cy,cz), r = rmin  */
  double cx, cy, cz;
  double stretch_z;               /* tanh clustering in z (0 = uniform)      */
  double lz;                      /* z extent when stretched (else n*h)      */
  int cylinder;                   /* solid cylinder along y                  */
  double cyl_d;                   /* cylinder diameter / x extent            */
} psb_case;

/* Writes the seven compact arrays, b, dV, the off mask (may be NULL) and the
 * per-direction cell sizes dxc[nx], dyc[ny], dzc[nz] (may be NULL).         */
int psb_generate(const psb_case *cs, double *c, double *w, double *e,
                 double *s, double *n, double *b_, double *t, double *rhs,
                 unsigned char *off, double *dxc, double *dyc, double *dzc);

/* ------------------------------------------------------------------------
 * The reference's own multigrid solve: AC::cycle (Src/Solver/Additive/
 * additive.h:30-95 vcycle / fcycle / wcycle / fullcycle, additive_cycle.cpp:
 * 4-99, additive_recursive.cpp:4-163) on the GPU in FP64: its levels
 * (smallest-prime-factor coarsening, domain_coarsen.cpp), its Galerkin
 * operators, CG smoothing with the diagonal preconditioner on every level,
 * restriction with mean removal, interpolation, the V, F and W cycles, the
 * full-cycle start, and its stopping, divergence, min-cycle and linf
 * controls (additive_cycle_functions.cpp:12-87, additive.h:97-142).
 * Cycle::flex (additive_recursive.cpp:161-200, AC::flexcycle) with a bound:
 * its recursion climbs back only when a level's smoother meets its ResRat,
 * and where one cannot, the reference's own build goes down and up between
 * two levels without end; here a flex cycle stops after max_visits level
 * visits, and the run ends as after a divergence (PSB_DIVERGED, x the
 * previous cycle's phi).  Where the reference's recursion ends, the result
 * is the same.
 * ------------------------------------------------------------------------ */
enum { PSB_CYCLE_NONE = -1, PSB_CYCLE_V = 0, PSB_CYCLE_F = 1, PSB_CYCLE_W = 2,
       PSB_CYCLE_FLEX = 3 };

typedef struct {
  int struct_size;    /* sizeof(psb_cycle), set by psb_cycle_defaults; other
                         values are PSB_EINVAL                               */
  double res_tol;     /* ResTol toler (< 0: off, as vcycle(ResRat))          */
  double res_rat;     /* ResRat factor (< 0: off, as vcycle(ResTol))         */
  int max_cycles;     /* AC::max_cycles (20, additive.cpp:11)                */
  int min_cycles;     /* AC::min_cycles (0)                                  */
  int max_iter[3];    /* AC::max_iter_array: CG iterations before the coarse
                         correction / re-smoothing between the two descents
                         of F and W / after it (20, 20, 20)                  */
  double dt;          /* time step: residual = sqrt(dot_voldiv_avg) * dt
                         (additive_residual.cpp:19-29); smoothers use 1/dt
                         as their scale (additive_cycle_functions.cpp:94)   */
  int loop;           /* the loop's cycle: PSB_CYCLE_V (default), _F, _W,
                         _FLEX                                              */
  int init;           /* PSB_CYCLE_NONE (default), or the shape of a full
                         cycle run before the loop (fullcycle(loop) uses V) */
  int stop_if_div;    /* AC::stop_if_diverging (1)                           */
  int priority_min_cyc; /* AC::priority_min_cycle (1)                        */
  int use_linf;       /* AC::use_linf_error (0): the stop tests, res0, res
                         and the history use max |r / dV| * dt              */
  double smooth_res_rat; /* AC::resrat: ResRat of the level solves (1e-18)   */
  double smooth_res_tol; /* AC::restol: ResTol of the level and coarsest
                            solves (1e-18)                                   */
  int max_stale[3];   /* AC::stale_iter_array: stalecount of the pre-, re-
                         and post-smoothing CG solves (-1, -1, -1 = off,
                         additive.cpp:19; at most 64); the coarsest solve
                         runs without one (additive_recursive.cpp:27)       */
  int max_visits;     /* PSB_CYCLE_FLEX: level visits per cycle before the
                         run is ended (1000; no counterpart in the
                         reference, whose flex recursion is unbounded)      */
  int level_solver;   /* the Krylov solver of every level and the coarsest:
                         PSB_CG (default) or PSB_BICGS, both with the
                         diagonal preconditioner -- the equation's solver
                         that AC::call_smoother / call_solver use
                         (additive_cycle_functions.cpp:89-107)            */
} psb_cycle;

void psb_cycle_defaults(psb_cycle *c);

/* x (in: initial guess, out: solution), b: host arrays in the context's
 * layout.  res->iterations = cycles, res->converged = the bool cycle()
 * returns; PSB_DIVERGED when a cycle did not reduce the residual (x then
 * holds the previous cycle's phi, as the reference restores it).
 * psb_history gives the residual before the first cycle, after the
 * full-cycle start (when init is set) and after every cycle.  A
 * non-symmetric matrix (7 arrays) keeps 7 arrays on every level, as the
 * reference's coarsen_system forms them (additive_coarsen_system.cpp).   */
int psb_ac_vcycle(psb_ctx *ctx, double *x, const double *b, const psb_cycle *c,
                  psb_result *res);
int psb_ac_vcycle_device(psb_ctx *ctx, double *d_x, const double *d_b,
                         const psb_cycle *c, psb_result *res);

/* Seeded batch of cases (dims from {32, 48, 64, 96, 128} per direction,
 * 1-8 bubbles of radius 0.05-0.2 of the shortest side, density ratio
 * 10-1000, own seeds); fills out[0..nsys).                                  */
int psb_batch_cases(unsigned long long seed, int nsys, psb_case *out);

/* ------------------------------------------------------------------------
 * Batched solves.  No counterpart in the reference, whose Linear::solve
 * takes one system: many independent pressure systems of different grid
 * sizes, each set up through its own context, solved together -- every
 * phase of an iteration is one launch over all systems, the iteration
 * sequence is one captured graph, each system stops on its own criteria.
 * The method is that of psb_solve with PSB_PREC_AC, PSB_CG, PSB_MIXED and
 * inner_adapt = 2 (FP32 AC-PCG with FP64 reliable updates); the coarsening
 * continues while a direction stays >= 4 cells.
 * ------------------------------------------------------------------------ */
typedef struct psb_batch psb_batch;

int psb_batch_create(psb_batch **bt, int nsys, const psb_layout *layouts);
int psb_batch_destroy(psb_batch *bt);

/* The context of system `sys` (owned by the batch): give it its geometry,
 * off mask and matrix with psb_set_geometry / psb_set_active /
 * psb_set_matrix.  The next batched solve rebuilds its plan if any system
 * changed.  NULL for a bad index.                                          */
psb_ctx *psb_batch_system(psb_batch *bt, int sys);

/* Tiles of the streaming kernels: 0 = chosen per (level, size class) when
 * the plan is built (default) by timing the Jacobi sweep of each class with
 * every candidate: the span tile with 1, 2, 4 or 8 cell vectors per thread,
 * and on levels of at least 2^15 cells whose rows hold whole 4-wide vectors
 * the marching tile (256 vectors of a (j, k) plane walking a slab of 4, 8,
 * 16 or 32 i planes).  With the environment variable PSB_TUNE_CACHE naming
 * a file, choices for this device are read from it (JSON) and new ones
 * written back.  > 0 = that many vectors per thread everywhere, the levels
 * of at least 2^15 cells marching with slabs that fill their segments.     */
int psb_batch_set_tuning(psb_batch *bt, int vpt);

/* The autotuner's choices: (level, log2 cells, choice) triples, choice > 0
 * vectors per thread of the span tile, < 0 minus the planes per slab of the
 * marching tile; returns their count.                                       */
int psb_batch_plan(psb_batch *bt, int *triples, int cap);

/* x[s] (in: initial guess, out: solution) and b[s] are host arrays in
 * system s's layout; res[s] its result.  One psb_stop for all.  Returns
 * PSB_OK when every system converged, else the first other status.          */
int psb_batch_solve(psb_batch *bt, double *const *x, const double *const *b,
                    const psb_stop *stop, psb_result *res);
/* Same on device-resident compact arrays (hipMalloc'd by the caller).      */
int psb_batch_solve_device(psb_batch *bt, double *const *d_x,
                           const double *const *d_b, const psb_stop *stop,
                           psb_result *res);

/* Benchmark hook: average device time (ms) of `reps` back-to-back launches
 * of one kernel on the context's stream, bracketed by HIP events.  kernel:
 * 0 FP64 SpMV, 1 FP32 SpMV, 2 FP32 damped Jacobi sweep on the fine level as
 * the last mixed-precision solve set it up (flux form; FP16 matrix copy with
 * ac_store = 1), 3 plain FP32 copy of the grid (reference stream for
 * hardware byte counters), 4 the fine level's two post-sweeps in one pass
 * (ac_fuse = 1 form), 5 the same two sweeps as two kernels, 6 the FP32 CG
 * p/q step of the mixed multigrid-preconditioned solve on the fine level,
 * 7 the fine level's first post-sweep with the prolongation fused in, 8 its
 * last post-sweep with the rho partials of the flexible beta, 9 the FP32 CG
 * x/r step.
 * *bytes receives the algorithmic bytes one launch moves (for 4 and 5:
 * those of one sweep).                                                      */
int psb_time_kernel(psb_ctx *ctx, int kernel, int reps, double *avg_ms,
                    double *bytes);

/* Benchmark hook: one solve exactly as psb_solve_device runs it (same
 * kernels in the same order on the same data), but launched kernel by kernel
 * instead of as captured graphs, each kernel on the finest level bracketed
 * by HIP events on the context's stream.  Per kernel kind: launches, their
 * summed duration (ms) and the algorithmic bytes one launch moves (the
 * arrays it must read and write once: matrix in its stored precision,
 * vectors, 1/8 of a coarse array for a restriction or prolongation).
 * *solve_ms: the whole eager solve between two events; what the fine-level
 * kernels leave of it is the coarse levels, the one-block reductions and
 * launch gaps.  The gated kernels of the reliable updates count every
 * launch, most of which return at once. */
enum psb_kernel_kind {
  PSB_K_SWEEP = 0,    /* damped Jacobi sweep (pre-smoothing, later post-sweeps) */
  PSB_K_SWEEP_PROL,   /* first post-sweep, prolongation of the coarse correction fused */
  PSB_K_SWEEP_LAST,   /* last post-sweep of the finest level, the rho partials fused */
  PSB_K_RESTRICT,     /* residual block sums to level 1, its first sweep fused */
  PSB_K_CG_PQ,        /* p = z + beta p, q = A p, p.q */
  PSB_K_CG_XR,        /* x += alpha p, r -= alpha q, norms, next first sweep */
  PSB_K_RU_RESID,     /* FP64 true residual of a reliable update (gated) */
  PSB_K_RU_XUPDATE,   /* x += d of a reliable update (gated) */
  PSB_K_COUNT
};
typedef struct {
  int launches;
  double ms;      /* summed over the launches */
  double bytes;   /* algorithmic bytes of one launch */
} psb_kernel_stat;
int psb_profile_solve_device(psb_ctx *ctx, double *d_x, const double *d_b,
                             const psb_stop *stop, psb_result *res,
                             psb_kernel_stat *stats, double *solve_ms);

#ifdef __cplusplus
}
#endif
#endif
