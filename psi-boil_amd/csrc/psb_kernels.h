// psb_kernels.h -- the hand-written gfx950 kernels of the pressure solver.
//
// Every streaming kernel reads the interior arrays once (HBM-bound; no
// MFMA): a block of 256 threads walks a contiguous run of V-wide cell
// vectors (V = 16 bytes / sizeof(T)), neighbours come through L1/L2, and the
// XCD-aware block remap keeps a slab's i+-1 planes in one L2.  Reductions
// are two-pass and deterministic: each streaming kernel leaves one partial
// per block, k_finalize sums them in a fixed order in FP64 (for both
// precisions) and applies the scalar step (alpha, beta, stop test, ...).
// Kernels of an iteration start with `if (st->done) return;` (skip_done for
// the preconditioner's kernels), so a captured
// graph of several iterations stops doing work once the solver has finished.
#pragma once

#include "../../include/psb.h"
#include "psb_device.h"

namespace psb {

// cell metrics for dV = dxc[i]*dyc[j]*dzc[k] (domain_surfaces_volumes.cpp:62)
struct Metric {
  const double *dxc, *dyc, *dzc;  // NULL: unit cells
  __device__ __forceinline__ double dV(int i, int j, int k) const {
    if (!dxc) return 1.0;
    return dxc[i] * dyc[j] * dzc[k];
  }
};

struct Work {
  double *partial;  // kMaxBlocks * 4 doubles
};

// contiguous run of V-vectors per block of the span; P, i, j, k of the
// first cell.  PSB_TILE_BEGIN spans the whole grid.
#define PSB_SPAN_BEGIN(g, V, sp)                                              \
  const int nvec_ = (g).N / (V);                                              \
  const int bl_ = (sp).bl;                                                    \
  const int per_ = (nvec_ + (sp).nb - 1) / (sp).nb;                           \
  const int q0_ = bl_ * per_;                                                 \
  const int q1_ = min(q0_ + per_, nvec_);                                     \
  for (int q_ = q0_ + (int)threadIdx.x; q_ < q1_; q_ += kBlock) {             \
    const int P = q_ * (V);                                                   \
    int i, j, k;                                                              \
    (g).ijk(P, i, j, k);
// kernels that run on the small levels (a few thousand cells, a handful of
// waves on the chip): let the compiler hold every load of a cell vector in
// flight at once rather than keep registers for occupancy those launches
// never reach
#define PSB_SMALL __attribute__((amdgpu_waves_per_eu(1, 2)))
#define PSB_TILE_BEGIN(g, V)                                                  \
  const Span tsp_ = grid_span();                                              \
  PSB_SPAN_BEGIN(g, V, tsp_)
#define PSB_TILE_END }

// ---------------------------------------------------------------------------
// y = A x   (Scalar::operator=(A_x), scalar_operators.cpp:49-72)
// ---------------------------------------------------------------------------
template <typename T, int V, bool SYM>
__global__ __launch_bounds__(kBlock) void k_spmv(Grid g, DMat<T> A, const T *x, T *y) {
  ArrX<T, V> X{x};
  PSB_TILE_BEGIN(g, V)
  Gath<T, V> G = gather<T, V, SYM>(g, A, X, P, i, j, k);
  stv<T, V>(y + P, apply_A<T, V>(G));
  PSB_TILE_END
}

// y = b - A x   (Scalar::operator=(y_m_A_x), scalar_operators.cpp:22-46)
template <typename T, int V, bool SYM>
__global__ __launch_bounds__(kBlock) void k_resid(Grid g, DMat<T> A, const T *b, const T *x,
                                                  T *y) {
  ArrX<T, V> X{x};
  PSB_TILE_BEGIN(g, V)
  Gath<T, V> G = gather<T, V, SYM>(g, A, X, P, i, j, k);
  stv<T, V>(y + P, resid<T, V>(ldv<T, V>(b + P), G));
  PSB_TILE_END
}

// ---------------------------------------------------------------------------
// Krylov start: r = b - A x with its norms, and per MODE
//   0 diagonal CG: rho = r.(r/c)        (cg.cpp:47-80)
//   1 preconditioned CG: e1 = omega r / c into aux
//   2 BiCGS: r~ = r into aux, rho = r~.r (bicgs.cpp:53-84)
// partials: sum r^2/dV^2, sum r^2, rho
// ---------------------------------------------------------------------------
template <typename T, int V, bool SYM, int MODE>
__global__ __launch_bounds__(kBlock) void k_start(Grid g, DMat<T> A, Metric m, const T *b,
                                                  const T *x, T *r, T *aux, T omega, Work w) {
  ArrX<T, V> X{x};
  double acc[3] = {0.0, 0.0, 0.0};
  PSB_TILE_BEGIN(g, V)
  Gath<T, V> G = gather<T, V, SYM>(g, A, X, P, i, j, k);
  VecN<T, V> R = resid<T, V>(ldv<T, V>(b + P), G), E;
  stv<T, V>(r + P, R);
#pragma unroll
  for (int q = 0; q < V; q++) {
    double rd = (double)R.v[q];
    double dv = m.dV(i, j, k + q);
    acc[0] += rd * rd / dv / dv;
    acc[1] += rd * rd;
    if (MODE == 0) {
      T z = R.v[q] * (T(1) / G.c.v[q]);
      acc[2] += (double)(R.v[q] * z);
    } else if (MODE == 1) {
      E.v[q] = omega * R.v[q] / G.c.v[q];
    } else {
      acc[2] += rd * rd;
    }
  }
  if (MODE == 1) stv<T, V>(aux + P, E);
  if (MODE == 2) stv<T, V>(aux + P, R);
  PSB_TILE_END
  block_partial<3>(acc, w.partial);
}

// ---------------------------------------------------------------------------
// CG step 1: p = z + beta p_old (z = r/c for the diagonal preconditioner,
// else the projected preconditioner output), q = A p; partial p.q
// (cg.cpp:75-108)
// ---------------------------------------------------------------------------
template <typename T, int V, bool SYM, bool DIAG, bool FLUX = false>
__device__ __forceinline__ void cg_pq_body(Span sp, const Grid &g, const DMat<T> &A,
                                           const T *r_or_z, const T *pold, T *pnew, T *q,
                                           const uint8_t *af, const DevState *st, Work w) {
  const T beta = (T)st->beta;
  double acc[1] = {0.0};
  PSB_SPAN_BEGIN(g, V, sp)
  Gath<T, V> G;
  if (DIAG) {
    DiagPX<T, V> X{r_or_z, A.c, pold, beta};
    G = gather<T, V, SYM>(g, A, X, P, i, j, k);
  } else {
    ZPX<T, V> X{r_or_z, pold, af, beta, (T)st->zmean};
    G = gather<T, V, SYM, FLUX>(g, A, X, P, i, j, k);
  }
  VecN<T, V> Q = apply_A<T, V>(G);
  stv<T, V>(pnew + P, G.x);
  stv<T, V>(q + P, Q);
#pragma unroll
  for (int u = 0; u < V; u++) acc[0] += (double)(G.x.v[u] * Q.v[u]);
  PSB_TILE_END
  block_partial<1>(acc, w.partial);
}

template <typename T, int V, bool SYM, bool DIAG, bool FLUX = false>
__global__ __launch_bounds__(kBlock) void k_cg_pq(Grid g, DMat<T> A, const T *r_or_z,
                                                  const T *pold, T *pnew, T *q,
                                                  const uint8_t *af, DevState *st, Work w) {
  if (st->done) return;
  cg_pq_body<T, V, SYM, DIAG, FLUX>(grid_span(), g, A, r_or_z, pold, pnew, q, af, st, w);
}

// ---------------------------------------------------------------------------
// CG step 2: x += alpha p, r -= alpha q (cg.cpp:114-115); partials of the
// residual norms and, in diagonal mode, rho_next = r.(r/c); preconditioned
// mode writes e1 = omega r / c, the first Jacobi sweep of the next V-cycle.
// ---------------------------------------------------------------------------
// one cell vector's update (X, R, Pp, Q, C its x, r, p, q, c)
#define PSB_CG_XR_CELL                                        \
  _Pragma("unroll") for (int u = 0; u < V; u++) {             \
    X.v[u] = X.v[u] + alpha * Pp.v[u];                        \
    R.v[u] = R.v[u] - alpha * Q.v[u];                         \
    double rd = (double)R.v[u];                               \
    if (vd) {                                                 \
      double dv = m.dV(i, j, k + u);                          \
      acc[0] += rd * rd / dv / dv;                            \
    }                                                         \
    acc[1] += rd * rd;                                        \
    if (DIAG) {                                               \
      T z = R.v[u] * (T(1) / C.v[u]);                         \
      acc[2] += (double)(R.v[u] * z);                         \
    } else {                                                  \
      E.v[u] = omega * R.v[u] / C.v[u];                       \
    }                                                         \
  }                                                           \
  stv<T, V>(x + P, X);                                        \
  stv<T, V>(r + P, R);                                        \
  if (!DIAG) stv<T, V>(e1 + P, E);

template <typename T, int V, bool DIAG>
__device__ __forceinline__ void cg_xr_body(Span sp, const Grid &g, const T *c, const Metric &m,
                                           const T *p, const T *q, T *x, T *r, T *e1, T omega,
                                           const DevState *st, Work w) {
  const T alpha = (T)st->alpha;
  // the reference norm's sum (dV loads, two FP64 divisions per cell) only
  // when the solve stops on it
  const bool vd = st->norm == 0;
  double acc[3] = {0.0, 0.0, 0.0};  // voldiv, r^2, rho_next
  if constexpr (sizeof(T) == 8) {
    // FP64: software-pipelined, the next vector's five loads issued before
    // the current one is computed (the span loop of PSB_SPAN_BEGIN; measured
    // at 256^3: 210 against 214 us; in FP32 slower, 112 against 108 us)
    const int nvec = g.N / V;
    const int per = (nvec + sp.nb - 1) / sp.nb;
    const int q0 = sp.bl * per, q1 = min(q0 + per, nvec);
    int qq = q0 + (int)threadIdx.x;
    const int Pf = (qq < q1 ? qq : 0) * V;
    VecN<T, V> Xn = ldv<T, V>(x + Pf), Rn = ldv<T, V>(r + Pf), Pn = ldv<T, V>(p + Pf),
               Qn = ldv<T, V>(q + Pf), Cn = ldv<T, V>(c + Pf);
    for (; qq < q1; qq += kBlock) {
      const int P = qq * V;
      int i, j, k;
      g.ijk(P, i, j, k);
      VecN<T, V> X = Xn, R = Rn, Pp = Pn, Q = Qn, C = Cn, E;
      const int Pnx = (qq + kBlock < q1 ? qq + kBlock : qq) * V;
      Xn = ldv<T, V>(x + Pnx);
      Rn = ldv<T, V>(r + Pnx);
      Pn = ldv<T, V>(p + Pnx);
      Qn = ldv<T, V>(q + Pnx);
      Cn = ldv<T, V>(c + Pnx);
      PSB_CG_XR_CELL
    }
  } else {
    PSB_SPAN_BEGIN(g, V, sp)
    VecN<T, V> X = ldv<T, V>(x + P), R = ldv<T, V>(r + P), Pp = ldv<T, V>(p + P),
               Q = ldv<T, V>(q + P), C = ldv<T, V>(c + P), E;
    PSB_CG_XR_CELL
    PSB_TILE_END
  }
  block_partial<3>(acc, w.partial);
}
#undef PSB_CG_XR_CELL

template <typename T, int V, bool DIAG>
__global__ __launch_bounds__(kBlock) void k_cg_xr(Grid g, const T *c, Metric m, const T *p,
                                                  const T *q, T *x, T *r, T *e1, T omega,
                                                  DevState *st, Work w) {
  if (st->done) return;
  cg_xr_body<T, V, DIAG>(grid_span(), g, c, m, p, q, x, r, e1, omega, st, w);
}

// ---------------------------------------------------------------------------
// red-black symmetric Gauss-Seidel preconditioner, one colour in place: on
// the cells with (i + j + k) % 2 == colour, z_P += (sum a_nb z_nb) / c_P.
// Starting from z = r / c, colour 1 and then colour 0 give
// z = (D - L)^-1 D (D - U)^-1 r in red-black order (even cells first).  The
// lanes of the other colour are stored back unchanged; every neighbour an
// updated cell reads is of that other colour, so no value read is one this
// launch changes.  RHO (the second colour, z then final on every cell):
// partials of r.z for FIN_RHO (sum z, sum r and the count left 0: no mean
// is removed, as in the reference's CG).
// ---------------------------------------------------------------------------
template <typename T, int V, bool SYM, bool RHO>
__global__ __launch_bounds__(kBlock) void k_rbsgs(Grid g, DMat<T> A, T *z, const T *r,
                                                  int colour, DevState *st, Work w) {
  if (st->done) return;
  ArrX<T, V> X{z};
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  PSB_TILE_BEGIN(g, V)
  Gath<T, V> G = gather<T, V, SYM>(g, A, X, P, i, j, k);
  VecN<T, V> E;
#pragma unroll
  for (int u = 0; u < V; u++) {
    const bool act = ((i + j + k + u) & 1) == colour;
    E.v[u] = act ? G.x.v[u] + ((G.s1.v[u] + G.s2.v[u]) + G.s3.v[u]) / G.c.v[u] : G.x.v[u];
  }
  if (RHO) {
    VecN<T, V> R = ldv<T, V>(r + P);
#pragma unroll
    for (int u = 0; u < V; u++) acc[0] += (double)R.v[u] * (double)E.v[u];
  }
  stv<T, V>(z + P, E);
  PSB_TILE_END
  if (RHO) block_partial<4>(acc, w.partial);
}

// ---------------------------------------------------------------------------
// damped Jacobi sweep  e_out = e + omega (r - A e) / c, the operand either e
// or e + cfac P eH (prolongation fused into the first post-sweep).
// RHO (last post-sweep on level 0): kRhoParts partials, r.z, sum z, sum r,
// active cells, from which k_finalize forms rho = r.(z - mean z) (the
// constants span the null space of the Neumann operator, and the projection
// keeps the search directions out of it: in FP32 a drift along it eats the
// digits of the correction), and with qf (the flexible beta) z.q and the sum
// of q over the active cells, q = A p of the last search direction.
// ---------------------------------------------------------------------------
constexpr int kRhoParts = 6;

template <typename T>
__device__ __forceinline__ void rho_acc(double (&acc)[kRhoParts], T r, T z, const T *qf, int P,
                                        bool act) {
  if (!act) return;
  acc[0] += (double)r * (double)z;
  acc[1] += (double)z;
  acc[2] += (double)r;
  acc[3] += 1.0;
  if (qf) {
    const double q = (double)qf[P];
    acc[4] += (double)z * q;
    acc[5] += q;
  }
}

// the same with q already loaded (hasq: the flexible beta's terms)
template <typename T>
__device__ __forceinline__ void rho_accq(double (&acc)[kRhoParts], T r, T z, bool hasq, T q,
                                         bool act) {
  if (!act) return;
  acc[0] += (double)r * (double)z;
  acc[1] += (double)z;
  acc[2] += (double)r;
  acc[3] += 1.0;
  if (hasq) {
    acc[4] += (double)z * (double)q;
    acc[5] += (double)q;
  }
}

template <typename T, int V, bool SYM, bool PROL, bool RHO, typename M = T, bool FLUX = false>
__device__ __forceinline__ void jacobi_body(Span sp, const Grid &g, const DMat<M> &A,
                                            const T *r, const T *e, T *eout,
                                            const ProlX<T, V> &pr, T omega, const uint8_t *af,
                                            Work w, T zscale = T(1), const T *qf = nullptr) {
  double acc[kRhoParts] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  PSB_SPAN_BEGIN(g, V, sp)
  Gath<T, V> G;
  if (PROL) {
    G = gather<T, V, SYM, FLUX>(g, A, pr, P, i, j, k);
  } else {
    ArrX<T, V> X{e};
    G = gather<T, V, SYM, FLUX>(g, A, X, P, i, j, k);
  }
  VecN<T, V> R = ldv<T, V>(r + P);
  VecN<T, V> t = resid<T, V>(R, G), E;
#pragma unroll
  for (int u = 0; u < V; u++) {
    E.v[u] = G.x.v[u] + omega * t.v[u] / G.c.v[u];
    if (RHO) {
      E.v[u] = zscale * E.v[u];
      rho_acc<T>(acc, R.v[u], E.v[u], qf, P + u, !af || af[P + u]);
    }
  }
  stv<T, V>(eout + P, E);
  PSB_TILE_END
  if (RHO) block_partial<kRhoParts>(acc, w.partial);
}

template <typename T, int V, bool SYM, bool PROL, bool RHO, typename M = T, bool FLUX = false>
__global__ __launch_bounds__(kBlock) PSB_SMALL void k_jacobi(Grid g, DMat<M> A, const T *r, const T *e,
                                                   T *eout, ProlX<T, V> pr, T omega,
                                                   const uint8_t *af, DevState *st, Work w,
                                                   T zscale, const T *qf) {
  if (skip_done(st, g.N)) return;
  jacobi_body<T, V, SYM, PROL, RHO, M, FLUX>(grid_span(), g, A, r, e, eout, pr, omega, af, w,
                                             zscale, qf);
}

// ---------------------------------------------------------------------------
// restriction for the preconditioner: rH = sum over the fine block of
// (r - A e)  (the residual sum of additive_restriction.cpp:12-135 without the
// reference's re-expansion; solid fine cells left out, the transpose of the
// masked prolongation), 0 on inactive coarse cells, and the coarse level's
// first sweep eH1 = omega rH / cH fused in.  A thread covers CI planes x CJ
// rows x V*KV cells along k of the fine grid (whole coarse cells); fine
// cells are summed in (i, j, k) order within each coarse cell.
// ---------------------------------------------------------------------------
template <typename T, int V, int KV, bool SYM, int CI, int CJ, int CK, typename M = T,
          bool FLUX = false>
__device__ __forceinline__ void restrict_body(Span sp, const Grid &g, const DMat<M> &A,
                                              const T *r, const T *e, const uint8_t *af,
                                              const Grid &gH, const T *cH, const uint8_t *afH,
                                              T *rH, T *eH1, T omega) {
  constexpr int KE = V * KV;        // fine cells along k per thread
  constexpr int NC = KE / CK;       // coarse cells per thread
  ArrX<T, V> X{e};
  const int nkv = g.nz / KE;
  const int nwork = gH.nx * gH.ny * nkv;
  for (int q = sp.bl * kBlock + threadIdx.x; q < nwork; q += sp.nb * kBlock) {
    const int IJ = q / nkv, kv = q - IJ * nkv;
    const int I = IJ / gH.ny, J = IJ - I * gH.ny;
    // the coarse cells' diagonal and flags loaded with the fine cells' data
    const int Q0 = (I * gH.ny + J) * gH.nz + kv * NC;
    T cq[NC];
    bool actH[NC];
#pragma unroll
    for (int u = 0; u < NC; u++) {
      cq[u] = cH[Q0 + u];
      actH[u] = flag_on(flag_word<1>(afH, Q0 + u, cH), 0);
    }
    T s[NC];
#pragma unroll
    for (int u = 0; u < NC; u++) s[u] = T(0);
#pragma unroll
    for (int a = 0; a < CI; a++) {
#pragma unroll
      for (int bb = 0; bb < CJ; bb++) {
        const int i = I * CI + a, j = J * CJ + bb;
        T t[KE];
#pragma unroll
        for (int h = 0; h < KV; h++) {
          const int k = kv * KE + h * V;
          const int P = (i * g.ny + j) * g.nz + k;
          Gath<T, V> G = gather<T, V, SYM, FLUX>(g, A, X, P, i, j, k);
          VecN<T, V> R = resid<T, V>(ldv<T, V>(r + P), G);
          const uint32_t fw = flag_word<V>(af, P, r);
#pragma unroll
          for (int u = 0; u < V; u++) t[h * V + u] = keep_if(flag_on(fw, u), R.v[u]);
        }
#pragma unroll
        for (int u = 0; u < NC; u++)
#pragma unroll
          for (int cc = 0; cc < CK; cc++) s[u] += t[u * CK + cc];
      }
    }
#pragma unroll
    for (int u = 0; u < NC; u++) {
      const T v = keep_if(actH[u], s[u]);
      rH[Q0 + u] = v;
      eH1[Q0 + u] = omega * v / cq[u];
    }
  }
}

template <typename T, int V, int KV, bool SYM, int CI, int CJ, int CK, typename M = T,
          bool FLUX = false>
__global__ __launch_bounds__(kBlock) PSB_SMALL void k_restrict(Grid g, DMat<M> A, const T *r, const T *e,
                                                     const uint8_t *af, Grid gH, const T *cH,
                                                     const uint8_t *afH, T *rH, T *eH1,
                                                     T omega, DevState *st) {
  if (skip_done(st, g.N)) return;
  restrict_body<T, V, KV, SYM, CI, CJ, CK, M, FLUX>(grid_span(), g, A, r, e, af, gH, cH, afH,
                                                    rH, eH1, omega);
}

// coarsest level: z = G r, G the dense projected pseudo-inverse (FP64).
// One wave per row; r is read in T and promoted.
template <typename T>
__device__ __forceinline__ void gemv_body(int bl, int m, const double *Gm, const T *r, T *z) {
  const int lane = threadIdx.x & 63;
  const int row = bl * kWaves + (threadIdx.x >> 6);
  if (row >= m) return;
  const double *g = Gm + (size_t)row * m;
  double s = 0.0;
  // unrolled so that a lane's loads leave together (same summation order)
#pragma unroll 8
  for (int c = lane; c < m; c += 64) s += g[c] * (double)r[c];
  s = wave_sum(s);
  if (lane == 0) z[row] = (T)s;
}

template <typename T>
__global__ __launch_bounds__(kBlock) PSB_SMALL void k_gemv(int m, const double *Gm, const T *r, T *z,
                                                 DevState *st) {
  if (skip_done(st, m)) return;
  gemv_body<T>(blockIdx.x, m, Gm, r, z);
}

// ---------------------------------------------------------------------------
// iterative refinement: r32 = (float)(b - A x) computed in FP64; partials
// of the FP64 norms and of the mean of r32 over the active cells
// ---------------------------------------------------------------------------
template <int V, bool SYM>
__global__ __launch_bounds__(kBlock) void k_ir_resid(Grid g, DMat<double> A, Metric m,
                                                     const double *b, const double *x,
                                                     float *r32, const uint8_t *af, Work w) {
  ArrX<double, V> X{x};
  double acc[4] = {0.0, 0.0, 0.0, 0.0};  // r64^2, r64^2/dV^2, sum r32, active cells
  PSB_TILE_BEGIN(g, V)
  Gath<double, V> G = gather<double, V, SYM>(g, A, X, P, i, j, k);
  VecN<double, V> R = resid<double, V>(ldv<double, V>(b + P), G);
#pragma unroll
  for (int u = 0; u < V; u++) {
    float rf = (float)R.v[u];
    r32[P + u] = rf;
    double dv = m.dV(i, j, k + u);
    acc[0] += R.v[u] * R.v[u];
    acc[1] += R.v[u] * R.v[u] / dv / dv;
    if (!af || af[P + u]) {
      acc[2] += (double)rf;
      acc[3] += 1.0;
    }
  }
  PSB_TILE_END
  block_partial<4>(acc, w.partial);
}

// start of the inner FP32 solve: the right-hand side loses its mean over
// the active cells (the part no correction can remove: it lies outside the
// range of the singular Neumann operator), e1 = omega r / c; partial ||r||^2
template <int V>
__global__ __launch_bounds__(kBlock) void k_center(Grid g, const float *c, const uint8_t *af,
                                                   float *r, float *e1, float omega,
                                                   const DevState *st, Work w) {
  const float mean = (float)st->mean[0];
  double acc[1] = {0.0};
  PSB_TILE_BEGIN(g, V)
  VecN<float, V> R = ldv<float, V>(r + P), Cc = ldv<float, V>(c + P), E;
#pragma unroll
  for (int u = 0; u < V; u++) {
    if (!af || af[P + u]) R.v[u] = R.v[u] - mean;
    E.v[u] = omega * R.v[u] / Cc.v[u];
    acc[0] += (double)R.v[u] * (double)R.v[u];
  }
  stv<float, V>(r + P, R);
  stv<float, V>(e1 + P, E);
  PSB_TILE_END
  block_partial<1>(acc, w.partial);
}

// ---------------------------------------------------------------------------
// reliable updates: when the FP32 CG's recursive residual has dropped by
// delta since the last refresh (st->update set by k_finalize), the
// accumulated FP32 correction moves into the FP64 solution and the true
// residual b - A x replaces the recursive one; CG then continues with its
// search direction (no restart).  Both kernels do nothing unless a refresh
// is pending (gated = true) or they start the solve (gated = false).
// ---------------------------------------------------------------------------
template <int V>
__device__ __forceinline__ void ru_xupdate_body(Span sp, const Grid &g, double *x, float *d) {
  PSB_SPAN_BEGIN(g, V, sp)
#pragma unroll
  for (int u = 0; u < V; u++) {
    x[P + u] = x[P + u] + (double)d[P + u];
    d[P + u] = 0.0f;
  }
  PSB_TILE_END
}

template <int V>
__global__ __launch_bounds__(kBlock) void k_ru_xupdate(Grid g, double *x, float *d,
                                                       const DevState *st, bool gated) {
  if (gated && !st->update) return;
  ru_xupdate_body<V>(grid_span(), g, x, d);
}

// r32 = (float)(b - A x) in FP64, e1 = omega r32 / c32; partials of the FP64
// norms (L2 and reference) and of the FP32 residual
template <int V, bool SYM>
__device__ __forceinline__ void ru_resid_body(Span sp, const Grid &g, const DMat<double> &A,
                                              const Metric &m, const double *b, const double *x,
                                              const float *c32, float *r32, float *e1,
                                              float omega, Work w, bool vd) {
  ArrX<double, V> X{x};
  // r64^2, r64^2/dV^2 (vd: the reference norm), r32^2, b^2
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  PSB_SPAN_BEGIN(g, V, sp)
  Gath<double, V> G = gather<double, V, SYM>(g, A, X, P, i, j, k);
  const VecN<double, V> Bv = ldv<double, V>(b + P);
  VecN<double, V> R = resid<double, V>(Bv, G);
  VecN<float, V> Cf = ldv<float, V>(c32 + P), Rf, E;
#pragma unroll
  for (int u = 0; u < V; u++) {
    Rf.v[u] = (float)R.v[u];
    E.v[u] = omega * Rf.v[u] / Cf.v[u];
    acc[0] += R.v[u] * R.v[u];
    if (vd) {
      double dv = m.dV(i, j, k + u);
      acc[1] += R.v[u] * R.v[u] / dv / dv;
    }
    acc[2] += (double)Rf.v[u] * (double)Rf.v[u];
    acc[3] += Bv.v[u] * Bv.v[u];
  }
  stv<float, V>(r32 + P, Rf);
  stv<float, V>(e1 + P, E);
  PSB_TILE_END
  block_partial<4>(acc, w.partial);
}

template <int V, bool SYM>
__global__ __launch_bounds__(kBlock) void k_ru_resid(Grid g, DMat<double> A, Metric m,
                                                     const double *b, const double *x,
                                                     const float *c32, float *r32, float *e1,
                                                     float omega, const DevState *st,
                                                     bool gated, Work w) {
  if (gated && !st->update) return;
  ru_resid_body<V, SYM>(grid_span(), g, A, m, b, x, c32, r32, e1, omega, w, st->norm == 0);
}

// x64 += (double) d32, d32 = 0
template <int V>
__global__ __launch_bounds__(kBlock) void k_ir_update(Grid g, double *x, float *d) {
  PSB_TILE_BEGIN(g, V)
#pragma unroll
  for (int u = 0; u < V; u++) {
    x[P + u] = x[P + u] + (double)d[P + u];
    d[P + u] = 0.0f;
  }
  PSB_TILE_END
}

// true relative residual ||b - A x|| / ||b|| (FP64 diagnostic after a solve)
template <int V, bool SYM>
__device__ __forceinline__ void relres_body(Span sp, const Grid &g, const DMat<double> &A,
                                            const double *b, const double *x, Work w) {
  ArrX<double, V> X{x};
  double acc[2] = {0.0, 0.0};
  PSB_SPAN_BEGIN(g, V, sp)
  Gath<double, V> G = gather<double, V, SYM>(g, A, X, P, i, j, k);
  VecN<double, V> B = ldv<double, V>(b + P), R = resid<double, V>(B, G);
#pragma unroll
  for (int u = 0; u < V; u++) {
    acc[0] += R.v[u] * R.v[u];
    acc[1] += B.v[u] * B.v[u];
  }
  PSB_TILE_END
  block_partial<2>(acc, w.partial);
}

template <int V, bool SYM>
__global__ __launch_bounds__(kBlock) void k_relres(Grid g, DMat<double> A, const double *b,
                                                   const double *x, Work w) {
  relres_body<V, SYM>(grid_span(), g, A, b, x, w);
}

// ---------------------------------------------------------------------------
// BiCGS (bicgs.cpp:79-216) in three passes per iteration.
// B1: p = r + beta (p_old - omega v_old) formed as the reference does
//     (p = r + beta p; p -= (beta omega) v), p^ = p / c (Diagonal), v = A p^;
//     partial r~.v
// ---------------------------------------------------------------------------
template <typename T, int V>
struct BiPX {
  const T *r, *p, *v, *c;
  T beta, bo;
  __device__ __forceinline__ T pnew(T rr, T pp, T vv) const {
    T t = rr + beta * pp;
    return t - bo * vv;
  }
  __device__ __forceinline__ VecN<T, V> vec(int P, int, int, int) const {
    VecN<T, V> R = ldv<T, V>(r + P), Pp = ldv<T, V>(p + P), Vv = ldv<T, V>(v + P),
               C = ldv<T, V>(c + P), o;
#pragma unroll
    for (int u = 0; u < V; u++) o.v[u] = pnew(R.v[u], Pp.v[u], Vv.v[u]) * (T(1) / C.v[u]);
    return o;
  }
  __device__ __forceinline__ T at(int Q, int, int, int) const {
    return pnew(r[Q], p[Q], v[Q]) * (T(1) / c[Q]);
  }
};

template <typename T, int V, bool SYM>
__global__ __launch_bounds__(kBlock) void k_bi1(Grid g, DMat<T> A, const T *r, const T *rt,
                                                const T *pold, const T *vold, T *pnew,
                                                T *vnew, DevState *st, Work w) {
  if (st->done) return;
  const bool first = st->iter == 0;
  const T beta = first ? T(0) : (T)st->beta;
  const T bo = first ? T(0) : (T)(st->beta * st->omega);
  BiPX<T, V> X{r, pold, vold, A.c, beta, bo};
  double acc[1] = {0.0};
  PSB_TILE_BEGIN(g, V)
  Gath<T, V> G = gather<T, V, SYM>(g, A, X, P, i, j, k);
  VecN<T, V> Vv = apply_A<T, V>(G), Rt = ldv<T, V>(rt + P), R = ldv<T, V>(r + P),
             Pp = ldv<T, V>(pold + P), Vo = ldv<T, V>(vold + P), Pn;
#pragma unroll
  for (int u = 0; u < V; u++) {
    Pn.v[u] = first ? R.v[u] : X.pnew(R.v[u], Pp.v[u], Vo.v[u]);
    acc[0] += (double)(Rt.v[u] * Vv.v[u]);
  }
  stv<T, V>(pnew + P, Pn);
  stv<T, V>(vnew + P, Vv);
  PSB_TILE_END
  block_partial<1>(acc, w.partial);
}

// B2: s = r - alpha v, s^ = s / c, t = A s^; partials s.s, t.s, t.t
template <typename T, int V>
struct BiSX {
  const T *r, *v, *c;
  T alpha;
  __device__ __forceinline__ T s1(T rr, T vv) const { return rr - alpha * vv; }
  __device__ __forceinline__ VecN<T, V> vec(int P, int, int, int) const {
    VecN<T, V> R = ldv<T, V>(r + P), Vv = ldv<T, V>(v + P), C = ldv<T, V>(c + P), o;
#pragma unroll
    for (int u = 0; u < V; u++) o.v[u] = s1(R.v[u], Vv.v[u]) * (T(1) / C.v[u]);
    return o;
  }
  __device__ __forceinline__ T at(int Q, int, int, int) const {
    return s1(r[Q], v[Q]) * (T(1) / c[Q]);
  }
};

template <typename T, int V, bool SYM>
__global__ __launch_bounds__(kBlock) void k_bi2(Grid g, DMat<T> A, const T *r, const T *v,
                                                T *s, T *t, DevState *st, Work w) {
  if (st->done) return;
  BiSX<T, V> X{r, v, A.c, (T)st->alpha};
  double acc[3] = {0.0, 0.0, 0.0};  // ss, ts, tt
  PSB_TILE_BEGIN(g, V)
  Gath<T, V> G = gather<T, V, SYM>(g, A, X, P, i, j, k);
  VecN<T, V> Tt = apply_A<T, V>(G), R = ldv<T, V>(r + P), Vv = ldv<T, V>(v + P), S;
#pragma unroll
  for (int u = 0; u < V; u++) {
    S.v[u] = X.s1(R.v[u], Vv.v[u]);
    acc[0] += (double)(S.v[u] * S.v[u]);
    acc[1] += (double)(Tt.v[u] * S.v[u]);
    acc[2] += (double)(Tt.v[u] * Tt.v[u]);
  }
  stv<T, V>(s + P, S);
  stv<T, V>(t + P, Tt);
  PSB_TILE_END
  block_partial<3>(acc, w.partial);
}

// B3: x += alpha p^; x += omega s^; r = s - omega t; partials of the norms
// and rho_next = r~.r.  With a small s (bicgs.cpp:134-138) only x += alpha p^.
template <typename T, int V>
__global__ __launch_bounds__(kBlock) void k_bi3(Grid g, const T *c, Metric m, const T *p,
                                                const T *s, const T *t, const T *rt, T *x,
                                                T *r, DevState *st, Work w) {
  if (st->done) return;
  const bool small_s = sqrt(st->rv_ss) < st->res_tol;
  const T alpha = (T)st->alpha, omega = (T)st->omega;
  double acc[3] = {0.0, 0.0, 0.0};  // voldiv, r^2, rho_next
  PSB_TILE_BEGIN(g, V)
  VecN<T, V> X = ldv<T, V>(x + P), Pp = ldv<T, V>(p + P), C = ldv<T, V>(c + P);
  if (small_s) {
#pragma unroll
    for (int u = 0; u < V; u++) X.v[u] = X.v[u] + alpha * (Pp.v[u] * (T(1) / C.v[u]));
    stv<T, V>(x + P, X);
  } else {
    VecN<T, V> S = ldv<T, V>(s + P), Tt = ldv<T, V>(t + P), Rt = ldv<T, V>(rt + P), R;
#pragma unroll
    for (int u = 0; u < V; u++) {
      T mc = T(1) / C.v[u];
      X.v[u] = X.v[u] + alpha * (Pp.v[u] * mc);
      X.v[u] = X.v[u] + omega * (S.v[u] * mc);
      R.v[u] = S.v[u] - omega * Tt.v[u];
      double rd = (double)R.v[u];
      double dv = m.dV(i, j, k + u);
      acc[0] += rd * rd / dv / dv;
      acc[1] += rd * rd;
      acc[2] += (double)(Rt.v[u] * R.v[u]);
    }
    stv<T, V>(x + P, X);
    stv<T, V>(r + P, R);
  }
  PSB_TILE_END
  block_partial<3>(acc, w.partial);
}

// ---------------------------------------------------------------------------
// BiCGS preconditioned by the multigrid cycle (not in the reference, whose
// BiCGS takes a diagonal or IC preconditioner): the steps of bicgs.cpp:79-229
// with prec->solve replaced by the cycle, applied to p and to s.  The passes
// around the two cycle applications; the scalar steps are FIN_BI1..3.
// A1: p = r + beta (p_old - omega v_old) (p = r at i = 0), as k_bi1 forms
//     it, and the level-0 first sweep e0 = omj p / c of the cycle on p
// ---------------------------------------------------------------------------
template <typename T, int V>
__global__ __launch_bounds__(kBlock) void k_abi_p(Grid g, const T *c, const T *r,
                                                  const T *pold, const T *vold, T *pnew, T *e0,
                                                  T omj, const DevState *st) {
  if (st->done) return;
  const bool first = st->iter == 0;
  const T beta = first ? T(0) : (T)st->beta;
  const T bo = first ? T(0) : (T)(st->beta * st->omega);
  PSB_TILE_BEGIN(g, V)
  const VecN<T, V> R = ldv<T, V>(r + P), Pp = ldv<T, V>(pold + P), Vv = ldv<T, V>(vold + P),
                   Cc = ldv<T, V>(c + P);
  VecN<T, V> Pn, E;
#pragma unroll
  for (int u = 0; u < V; u++) {
    const T t = R.v[u] + beta * Pp.v[u];
    Pn.v[u] = first ? R.v[u] : t - bo * Vv.v[u];
    E.v[u] = omj * Pn.v[u] / Cc.v[u];
  }
  stv<T, V>(pnew + P, Pn);
  if (e0) stv<T, V>(e0 + P, E);
  PSB_TILE_END
}

// A2: v = A z1 with z1 = M^-1 p (kept in zs for the x update), partial r~.v
template <typename T, int V, bool SYM>
__global__ __launch_bounds__(kBlock) void k_abi_v(Grid g, DMat<T> A, const T *z1, const T *rt,
                                                  T *v, T *zs, DevState *st, Work w) {
  if (st->done) return;
  const ArrX<T, V> X{z1};
  double acc[1] = {0.0};
  PSB_TILE_BEGIN(g, V)
  const VecN<T, V> Rt = ldv<T, V>(rt + P);
  const Gath<T, V> G = gather<T, V, SYM>(g, A, X, P, i, j, k);
  const VecN<T, V> Vv = apply_A<T, V>(G);
#pragma unroll
  for (int u = 0; u < V; u++) acc[0] += (double)(Rt.v[u] * Vv.v[u]);
  stv<T, V>(v + P, Vv);
  if (zs) stv<T, V>(zs + P, G.x);
  PSB_TILE_END
  block_partial<1>(acc, w.partial);
}

// A3: s = r - alpha v, and the level-0 first sweep e0 = omj s / c of the
// cycle on s
template <typename T, int V>
__global__ __launch_bounds__(kBlock) void k_abi_s(Grid g, const T *c, const T *r, const T *v,
                                                  T *s, T *e0, T omj, const DevState *st) {
  if (st->done) return;
  const T alpha = (T)st->alpha;
  PSB_TILE_BEGIN(g, V)
  const VecN<T, V> R = ldv<T, V>(r + P), Vv = ldv<T, V>(v + P), Cc = ldv<T, V>(c + P);
  VecN<T, V> S, E;
#pragma unroll
  for (int u = 0; u < V; u++) {
    S.v[u] = R.v[u] - alpha * Vv.v[u];
    E.v[u] = omj * S.v[u] / Cc.v[u];
  }
  stv<T, V>(s + P, S);
  if (e0) stv<T, V>(e0 + P, E);
  PSB_TILE_END
}

// A4: t = A z2 with z2 = M^-1 s; partials s.s, t.s, t.t (FIN_BI2's order)
template <typename T, int V, bool SYM>
__global__ __launch_bounds__(kBlock) void k_abi_t(Grid g, DMat<T> A, const T *z2, const T *s,
                                                  T *t, DevState *st, Work w) {
  if (st->done) return;
  const ArrX<T, V> X{z2};
  double acc[3] = {0.0, 0.0, 0.0};
  PSB_TILE_BEGIN(g, V)
  const VecN<T, V> S = ldv<T, V>(s + P);
  const Gath<T, V> G = gather<T, V, SYM>(g, A, X, P, i, j, k);
  const VecN<T, V> Tt = apply_A<T, V>(G);
#pragma unroll
  for (int u = 0; u < V; u++) {
    acc[0] += (double)(S.v[u] * S.v[u]);
    acc[1] += (double)(Tt.v[u] * S.v[u]);
    acc[2] += (double)(Tt.v[u] * Tt.v[u]);
  }
  stv<T, V>(t + P, Tt);
  PSB_TILE_END
  block_partial<3>(acc, w.partial);
}

// mixed form (FP32 V-cycle inside the FP64 BiCGS): x32 = (float)x and the
// FP32 level-0 first sweep e32 = omj x32 / c32; back: z64 = (double)z32
template <int V>
__global__ __launch_bounds__(kBlock) void k_abi_to32(Grid g, const double *x, const float *c32,
                                                     float *x32, float *e32, float omj,
                                                     const DevState *st) {
  if (st->done) return;
  PSB_TILE_BEGIN(g, V)
  const VecN<double, V> X = ldv<double, V>(x + P);
  const VecN<float, V> Cc = ldv<float, V>(c32 + P);
  VecN<float, V> Y, E;
#pragma unroll
  for (int u = 0; u < V; u++) {
    Y.v[u] = (float)X.v[u];
    E.v[u] = omj * Y.v[u] / Cc.v[u];
  }
  stv<float, V>(x32 + P, Y);
  stv<float, V>(e32 + P, E);
  PSB_TILE_END
}

template <int V>
__global__ __launch_bounds__(kBlock) void k_abi_to64(Grid g, const float *z32, double *z,
                                                     const DevState *st) {
  if (st->done) return;
  PSB_TILE_BEGIN(g, V)
  const VecN<float, V> Z = ldv<float, V>(z32 + P);
  VecN<double, V> Y;
#pragma unroll
  for (int u = 0; u < V; u++) Y.v[u] = (double)Z.v[u];
  stv<double, V>(z + P, Y);
  PSB_TILE_END
}

// A5: x += alpha z1; x += omega z2; r = s - omega t; partials of the norms
// and rho_next = r~.r; with a small s only x += alpha z1 (as k_bi3)
template <typename T, int V>
__global__ __launch_bounds__(kBlock) void k_abi_x(Grid g, Metric m, const T *z1, const T *z2,
                                                  const T *s, const T *t, const T *rt, T *x,
                                                  T *r, DevState *st, Work w) {
  if (st->done) return;
  const bool small_s = sqrt(st->rv_ss) < st->res_tol;
  const T alpha = (T)st->alpha, omega = (T)st->omega;
  double acc[3] = {0.0, 0.0, 0.0};  // voldiv, r^2, rho_next
  PSB_TILE_BEGIN(g, V)
  VecN<T, V> X = ldv<T, V>(x + P);
  const VecN<T, V> Z1 = ldv<T, V>(z1 + P);
  if (small_s) {
#pragma unroll
    for (int u = 0; u < V; u++) X.v[u] = X.v[u] + alpha * Z1.v[u];
    stv<T, V>(x + P, X);
  } else {
    const VecN<T, V> Z2 = ldv<T, V>(z2 + P), S = ldv<T, V>(s + P), Tt = ldv<T, V>(t + P),
                     Rt = ldv<T, V>(rt + P);
    VecN<T, V> R;
#pragma unroll
    for (int u = 0; u < V; u++) {
      X.v[u] = X.v[u] + alpha * Z1.v[u];
      X.v[u] = X.v[u] + omega * Z2.v[u];
      R.v[u] = S.v[u] - omega * Tt.v[u];
      const double rd = (double)R.v[u];
      const double dv = m.dV(i, j, k + u);
      acc[0] += rd * rd / dv / dv;
      acc[1] += rd * rd;
      acc[2] += (double)(Rt.v[u] * R.v[u]);
    }
    stv<T, V>(x + P, X);
    stv<T, V>(r + P, R);
  }
  PSB_TILE_END
  block_partial<3>(acc, w.partial);
}

// ---------------------------------------------------------------------------
// pass 2 of the reductions: one block sums the partials of the kernel
// before it (fixed order) and applies the scalar step
// ---------------------------------------------------------------------------
enum FinOp {
  FIN_START_DIAG, FIN_START_AC, FIN_START_BICGS, FIN_PQ, FIN_XR_DIAG, FIN_XR_AC, FIN_RHO,
  FIN_XR_RU, FIN_IR_RESID, FIN_CENTER, FIN_RELRES, FIN_BI1, FIN_BI2, FIN_BI3, FIN_RU_START,
  FIN_RU, FIN_RU_END, FIN_ACMEAN, FIN_ACRES, FIN_RHO_FLEX
};

// stop logic shared by the Krylov loops (cg.cpp:121-160, bicgs.cpp:174-210),
// after iteration st->iter with the new residual
__device__ inline void krylov_stop(DevState *st, double res, bool small_omega) {
  const int i = st->iter;
  st->res = res;
  if (st->hist && i + 1 < st->hist_cap) st->hist[i + 1] = res;
  int done = 0, conv = 0, reason = 0;
  if (!(res == res) || res > 1e300) {
    done = 1; reason = 8;
  } else if (res < st->res_tol && i >= st->mini - 1) {
    done = 1; conv = 1; reason = 1;
  } else if (res < st->res0 * st->res_rat && i >= st->mini - 1) {
    done = 1; conv = 1; reason = 2;
  } else if (small_omega) {
    done = 1; reason = 6;
  } else if (st->stalecount > 0) {
    bool stale = true;
    for (int q = 0; q < st->stalecount; q++)
      if (res < st->stale[q]) { stale = false; break; }
    if (stale) {
      done = 1; reason = 3;
    } else {
      for (int q = 0; q + 1 < st->stalecount; q++) st->stale[q] = st->stale[q + 1];
      st->stale[st->stalecount - 1] = res;
    }
  }
  if (done) {
    st->report_iters = i + 1;
  } else if (i + 1 >= st->maxit) {
    done = 1;
    st->report_iters = st->maxit + 1;
  }
  st->converged = conv;
  st->reason = reason;
  st->iter = i + 1;
  st->done = done;
}

__device__ inline double ref_norm(const DevState *st, double sum_voldiv) {
  // sqrt(dot_voldiv_avg)/scale with dot_voldiv_avg = d/ntot/ntot
  // (scalar.h:317-326, cg.cpp:121)
  double d = sum_voldiv / st->ntot;
  d = d / st->ntot;
  return sqrt(d) / st->scale;
}

__device__ inline void krylov_init(DevState *st, double res, double rsq, double rho) {
  st->res0 = res;
  st->res = res;
  st->rsq0 = rsq;
  st->rsq = rsq;
  st->rho = rho;
  st->beta = 0.0;
  st->zmean = 0.0;
  st->iter = 0;
  st->report_iters = 0;
  st->converged = 0;
  st->reason = 0;
  st->done = 0;
  if (st->hist && st->hist_cap > 0) st->hist[0] = res;
  for (int q = 0; q < st->stalecount && q < kStaleMax; q++) st->stale[q] = 1.0e24;
  if (st->stalecount > 0) st->stale[st->stalecount - 1] = res;
}

// steps of a finished loop (or a refresh not pending) do nothing
__device__ inline bool fin_skip(int op, const DevState *st) {
  switch (op) {
    case FIN_PQ: case FIN_XR_DIAG: case FIN_XR_AC: case FIN_RHO: case FIN_RHO_FLEX:
    case FIN_XR_RU:
    case FIN_BI1: case FIN_BI2: case FIN_BI3:
      return st->done != 0;
    case FIN_RU:
      return st->done || !st->update;
    default:
      return false;
  }
}

// sum of the partials of blocks [b0, b0 + nblk) in a fixed order (all
// threads of the block receive the totals)
template <int NR>
__device__ inline void fin_sum(int b0, int nblk, const double *partial, double (&acc)[NR]) {
  __shared__ double sm[kWaves][NR];
#pragma unroll
  for (int r = 0; r < NR; r++) acc[r] = 0.0;
  for (int q = threadIdx.x; q < nblk; q += kBlock) {
#pragma unroll
    for (int r = 0; r < NR; r++) acc[r] += partial[(size_t)(b0 + q) * NR + r];
  }
  block_sum<NR>(acc, sm);
}

// the scalar step of op on the totals (one thread)
__device__ inline void fin_apply(int op, const double *tot, DevState *st, DevState *st2,
                                 double *out) {
  switch (op) {
    case FIN_START_DIAG:
    case FIN_START_AC:
    case FIN_START_BICGS: {
      const double res = st->norm == 0 ? ref_norm(st, tot[0]) : sqrt(tot[1]);
      krylov_init(st, res, tot[1], tot[2]);
      if (res < st->res_tol) {  // cg.cpp:64, bicgs.cpp:70 early return
        st->done = 1;
        st->converged = 1;
        st->reason = 7;
      }
      if (!(res == res)) {
        st->done = 1;
        st->reason = 8;
      }
      if (op == FIN_START_BICGS && !st->done && sqrt(fabs(tot[2])) < st->res_tol) {
        st->done = 1;  // bicgs.cpp:89 at i = 0
        st->reason = 4;
        st->report_iters = 1;
      }
      break;
    }
    case FIN_PQ:
      st->pq = tot[0];
      st->alpha = st->rho / tot[0];  // cg.cpp:107-108
      break;
    case FIN_XR_DIAG:
    case FIN_XR_AC: {
      const double res = st->norm == 0 ? ref_norm(st, tot[0]) : sqrt(tot[1]);
      st->rsq = tot[1];
      krylov_stop(st, res, false);
      if (op == FIN_XR_DIAG && !st->done) {
        st->beta = tot[2] / st->rho;  // cg.cpp:93
        st->rho = tot[2];
      }
      break;
    }
    case FIN_XR_RU: {  // FP32 step inside the refinement: refresh or stop
      const double res = sqrt(tot[1]);
      const int i = st->iter;
      st->rsq = tot[1];
      st->res = res;
      st->iter = i + 1;
      if (!(res == res) || res > 1e300) {
        st->done = 1;
        st->reason = 8;
        st->report_iters = i + 1;
      } else if (res < st->rlast * st->delta || res < sqrt(st->rsq0) * st->res_rat) {
        st->update = 1;
      } else if (i + 1 >= st->maxit) {
        st->done = 1;
        st->report_iters = st->maxit + 1;
      }
      break;
    }
    case FIN_RU_START:
    case FIN_RU: {  // true residual after a refresh (or at the start)
      const double rl2 = sqrt(tot[0]);
      const double res = st->norm == 0 ? ref_norm(st, tot[1]) : rl2;
      if (op == FIN_RU_START) {
        krylov_init(st, res, tot[0], 0.0);
        st->outer = 0;
      } else {
        st->outer += 1;
        if (st->hist && st->outer < st->hist_cap) st->hist[st->outer] = res;
      }
      st->res_true = res;
      st->rl2_true = rl2;
      if (op == FIN_RU_START) st->bsq = tot[3];
      st->rlast = rl2;
      st->update = 0;
      st->res = sqrt(tot[2]);
      if (!(res == res)) {
        st->done = 1;
        st->reason = 8;
      } else if (res < st->res_tol || res < st->res0 * st->res_rat || st->res0 == 0.0) {
        st->done = 1;
        st->converged = 1;
        st->reason = 2;
        st->report_iters = st->iter;
      } else if (st->iter >= st->maxit) {
        st->done = 1;
        st->report_iters = st->maxit + 1;
      }
      break;
    }
    case FIN_ACMEAN:  // mean of the restricted right-hand side over active cells
      st->mean[0] = tot[1] > 0.0 ? tot[0] / tot[1] : 0.0;
      break;
    case FIN_ACRES: {  // AC::residual: sqrt(dot_voldiv_avg(r, r)) * dt (scale = dt)
      double d = tot[0] / st->ntot;
      d = d / st->ntot;
      out[0] = sqrt(d) * st->scale;
      break;
    }
    case FIN_RU_END:  // true residual at exit without convergence
      st->res_true = st->norm == 0 ? ref_norm(st, tot[1]) : sqrt(tot[0]);
      st->rl2_true = sqrt(tot[0]);
      break;
    case FIN_RHO:
    case FIN_RHO_FLEX: {
      const double m = !st->keepmean && tot[3] > 0.0 ? tot[1] / tot[3] : 0.0;
      const double rho = tot[0] - m * tot[2];
      st->zmean = m;
      if (st->iter == 0) {
        st->beta = 0.0;
      } else if (op == FIN_RHO) {
        st->beta = rho / st->rho;  // Fletcher-Reeves, cg.cpp:93
      } else {
        // flexible: beta = -z'.A p / p.A p with z' = z - m on the active
        // cells, so the new direction is A-orthogonal to the last one
        // whatever the preconditioner did (Notay, SISC 22 (2000) 1444-1460)
        const double zq = tot[4] - m * tot[5];
        st->beta = st->pq != 0.0 ? -zq / st->pq : 0.0;
      }
      st->rho = rho;
      break;
    }
    case FIN_IR_RESID:  // st: FP64 outer state, st2: FP32 inner state
      st->rsq = tot[0];
      st->res = st->norm == 0 ? ref_norm(st, tot[1]) : sqrt(tot[0]);
      st2->mean[0] = !st2->keepmean && tot[3] > 0.0 ? tot[2] / tot[3] : 0.0;
      break;
    case FIN_CENTER: {
      krylov_init(st, sqrt(tot[0]), tot[0], 0.0);
      if (tot[0] == 0.0) st->done = 1;
      break;
    }
    case FIN_RELRES:
      out[0] = tot[0];
      out[1] = tot[1];
      break;
    case FIN_BI1:
      st->rv = tot[0];
      st->alpha = st->rho / tot[0];  // bicgs.cpp:122-123
      break;
    case FIN_BI2:
      st->rv_ss = tot[0];
      st->rv_ts = tot[1];
      st->rv_tt = tot[2];
      st->omega = tot[1] / tot[2];  // bicgs.cpp:155-157
      break;
    case FIN_BI3: {
      if (sqrt(st->rv_ss) < st->res_tol) {  // bicgs.cpp:134-138
        st->done = 1;
        st->converged = 0;
        st->reason = 5;
        st->report_iters = st->iter + 1;
        break;
      }
      const double res = st->norm == 0 ? ref_norm(st, tot[0]) : sqrt(tot[1]);
      st->rsq = tot[1];
      const bool small_om = st->omega >= -st->res_tol && st->omega <= st->res_tol;
      krylov_stop(st, res, small_om);
      if (!st->done) {
        // top of the next iteration: rho = r~.r, breakdown test (bicgs.cpp:84-89)
        const double rho = tot[2];
        if (sqrt(fabs(rho)) < st->res_tol) {
          st->done = 1;
          st->reason = 4;
          st->report_iters = st->iter + 1;
        } else {
          st->beta = (rho / st->rho) * (st->alpha / st->omega);  // bicgs.cpp:102
          st->rho = rho;
        }
      }
      break;
    }
  }
}

template <int NR>
__global__ __launch_bounds__(kBlock) void k_finalize(int op, int nblk, const double *partial,
                                                     DevState *st, DevState *st2, double *out) {
  if (fin_skip(op, st)) return;
  double acc[NR];
  fin_sum<NR>(0, nblk, partial, acc);
  if (threadIdx.x == 0) fin_apply(op, acc, st, st2, out);
}

// ---------------------------------------------------------------------------
// K-cycle on a coarse level (two flexible-CG steps preconditioned by the
// cycle below; Y. Notay and P. S. Vassilevski, Numer. Linear Algebra Appl. 15
// (2008) 473-487).  The level's corrections are held divided by its FP16
// scale s, so its operator here is s A (A the FP32 or FP64 level matrix).
// k_kc_av: v = s A c, kc = c, partials c.v and c.r
// ---------------------------------------------------------------------------
template <typename T, int V, bool FLUX>
__device__ __forceinline__ void kc_av_body(Span sp, const Grid &g, const DMat<T> &A, T scale,
                                           const T *c, const T *r, T *kc, T *v, Work w) {
  ArrX<T, V> X{c};
  double acc[2] = {0.0, 0.0};
  PSB_SPAN_BEGIN(g, V, sp)
  Gath<T, V> G = gather<T, V, true, FLUX>(g, A, X, P, i, j, k);
  VecN<T, V> Av = apply_A<T, V>(G), R = ldv<T, V>(r + P);
#pragma unroll
  for (int u = 0; u < V; u++) {
    Av.v[u] = scale * Av.v[u];
    acc[0] += (double)G.x.v[u] * (double)Av.v[u];
    acc[1] += (double)G.x.v[u] * (double)R.v[u];
  }
  stv<T, V>(v + P, Av);
  stv<T, V>(kc + P, G.x);
  PSB_TILE_END
  block_partial<2>(acc, w.partial);
}

template <typename T, int V, bool FLUX>
__global__ __launch_bounds__(kBlock) PSB_SMALL void k_kc_av(Grid g, DMat<T> A, T scale, const T *c,
                                                  const T *r, T *kc, T *v, const DevState *st,
                                                  Work w) {
  if (skip_done(st, g.N)) return;
  kc_av_body<T, V, FLUX>(grid_span(), g, A, scale, c, r, kc, v, w);
}

// r <- r - alpha1 v, and the next cycle's first sweep e0 = omega r / c.  Every
// block sums k_kc_av's nprev partials itself (in k_finalize's order; a few
// hundred at most on these levels) instead of a finalize launch: rho1 =
// c.v, alpha1 = c.r / rho1, written to kst[0..1] by the span's block 0
template <typename T, int V>
__device__ __forceinline__ void kc_r_body(Span sp, const Grid &g, const T *c, const T *v, T *r,
                                          T *e0, T omega, double *kst, int nprev,
                                          const double *partial) {
  __shared__ double a1s;
  double tot[2];
  fin_sum<2>(0, nprev, partial, tot);
  if (threadIdx.x == 0) {
    a1s = tot[0] > 0.0 ? tot[1] / tot[0] : 0.0;
    if (sp.bl == 0) {
      kst[0] = tot[0];
      kst[1] = a1s;
    }
  }
  __syncthreads();
  const T a1 = (T)a1s;
  PSB_SPAN_BEGIN(g, V, sp)
  VecN<T, V> R = ldv<T, V>(r + P), Vv = ldv<T, V>(v + P), Cc = ldv<T, V>(c + P), E;
#pragma unroll
  for (int u = 0; u < V; u++) {
    R.v[u] = R.v[u] - a1 * Vv.v[u];
    E.v[u] = omega * R.v[u] / Cc.v[u];
  }
  stv<T, V>(r + P, R);
  stv<T, V>(e0 + P, E);
  PSB_TILE_END
}

template <typename T, int V>
__global__ __launch_bounds__(kBlock) PSB_SMALL void k_kc_r(Grid g, const T *c, const T *v, T *r, T *e0,
                                                 T omega, double *kst, int nprev,
                                                 const double *partial, const DevState *st) {
  if (skip_done(st, g.N)) return;
  kc_r_body<T, V>(grid_span(), g, c, v, r, e0, omega, kst, nprev, partial);
}

// k_kc_r fused with the second pre-sweep of the cycle that follows it (a
// cycle with nu1 >= 2): rt = r - alpha1 v (to rt, r is read by the
// neighbours), and e1 = e0 + omega (rt - A e0) / c with the first sweep
// e0 = omega0 rt / c0 (as k_kc_r writes it) formed at every point the stencil
// reads.  The same arithmetic as k_kc_r followed by k_jacobi, one launch
// fewer.
template <typename T, int V>
struct KrX {
  const T *r, *v, *c;
  T a1, om0;
  __device__ __forceinline__ T one(T rr, T vv, T cc) const { return om0 * (rr - a1 * vv) / cc; }
  __device__ __forceinline__ VecN<T, V> vec(int P, int, int, int) const {
    const VecN<T, V> R = ldv<T, V>(r + P), Vv = ldv<T, V>(v + P), C = ldv<T, V>(c + P);
    VecN<T, V> o;
#pragma unroll
    for (int u = 0; u < V; u++) o.v[u] = one(R.v[u], Vv.v[u], C.v[u]);
    return o;
  }
  __device__ __forceinline__ T at(int Q, int, int, int) const { return one(r[Q], v[Q], c[Q]); }
};

template <typename T, int V, typename M, bool FLUX>
__global__ __launch_bounds__(kBlock) PSB_SMALL void k_kc_r_sweep(Grid g, DMat<M> A, const T *c,
                                                       const T *v, const T *r, T *rt, T *e1,
                                                       T om0, T omega, double *kst, int nprev,
                                                       const double *partial,
                                                       const DevState *st) {
  if (skip_done(st, g.N)) return;
  __shared__ double a1s;
  double tot[2];
  fin_sum<2>(0, nprev, partial, tot);
  if (threadIdx.x == 0) {
    a1s = tot[0] > 0.0 ? tot[1] / tot[0] : 0.0;
    if (blockIdx.x == 0) {
      kst[0] = tot[0];
      kst[1] = a1s;
    }
  }
  __syncthreads();
  const T a1 = (T)a1s;
  const KrX<T, V> X{r, v, c, a1, om0};
  PSB_TILE_BEGIN(g, V)
  const VecN<T, V> R = ldv<T, V>(r + P), Vv = ldv<T, V>(v + P);
  const Gath<T, V> G = gather<T, V, true, FLUX>(g, A, X, P, i, j, k);
  VecN<T, V> Rt, E;
#pragma unroll
  for (int u = 0; u < V; u++) Rt.v[u] = R.v[u] - a1 * Vv.v[u];
  const VecN<T, V> t = resid<T, V>(Rt, G);
#pragma unroll
  for (int u = 0; u < V; u++) E.v[u] = G.x.v[u] + omega * t.v[u] / G.c.v[u];
  stv<T, V>(rt + P, Rt);
  stv<T, V>(e1 + P, E);
  PSB_TILE_END
}

// partials d.v, d.(s A d), d.r
template <typename T, int V, bool FLUX>
__device__ __forceinline__ void kc_dw_body(Span sp, const Grid &g, const DMat<T> &A, T scale,
                                           const T *d, const T *v, const T *r, Work w) {
  ArrX<T, V> X{d};
  double acc[3] = {0.0, 0.0, 0.0};
  PSB_SPAN_BEGIN(g, V, sp)
  Gath<T, V> G = gather<T, V, true, FLUX>(g, A, X, P, i, j, k);
  VecN<T, V> Ad = apply_A<T, V>(G), Vv = ldv<T, V>(v + P), R = ldv<T, V>(r + P);
#pragma unroll
  for (int u = 0; u < V; u++) {
    const double dd = (double)G.x.v[u];
    acc[0] += dd * (double)Vv.v[u];
    acc[1] += dd * (double)(scale * Ad.v[u]);
    acc[2] += dd * (double)R.v[u];
  }
  PSB_TILE_END
  block_partial<3>(acc, w.partial);
}

template <typename T, int V, bool FLUX>
__global__ __launch_bounds__(kBlock) PSB_SMALL void k_kc_dw(Grid g, DMat<T> A, T scale, const T *d,
                                                  const T *v, const T *r, const DevState *st,
                                                  Work w) {
  if (skip_done(st, g.N)) return;
  kc_dw_body<T, V, FLUX>(grid_span(), g, A, scale, d, v, r, w);
}

// out = a kc + alpha2 d, the coefficients from k_kc_dw's nprev partials
// (gamma = d.v, beta = d.s A d, delta = d.rt) summed by every block, with
// rho1 and alpha1 from kst: alpha2 = delta / (beta - gamma^2 / rho1),
// a = alpha1 - gamma alpha2 / rho1 (the span's block 0 keeps them in kst[2..3])
template <typename T, int V>
__device__ __forceinline__ void kc_comb_body(Span sp, const Grid &g, const T *kc, const T *d,
                                             T *out, double *kst, int nprev,
                                             const double *partial) {
  __shared__ double cs[2];
  // the first vector's loads (and rho1, alpha1) issued before the partial
  // sums, which the rest waits for anyway
  const int nvec = g.N / V, per = (nvec + sp.nb - 1) / sp.nb;
  const int q0 = sp.bl * per, q1 = min(q0 + per, nvec);
  const int Pf = (q0 + (int)threadIdx.x < q1 ? q0 + (int)threadIdx.x : 0) * V;
  VecN<T, V> Kc = ldv<T, V>(kc + Pf), D = ldv<T, V>(d + Pf);
  const double rho1 = kst[0], a1 = kst[1];
  double tot[3];
  fin_sum<3>(0, nprev, partial, tot);
  if (threadIdx.x == 0) {
    const double den = rho1 > 0.0 ? tot[1] - tot[0] * tot[0] / rho1 : 0.0;
    const double a2 = den > 0.0 ? tot[2] / den : 0.0;
    cs[0] = a1 - (rho1 > 0.0 ? tot[0] * a2 / rho1 : 0.0);
    cs[1] = a2;
    if (sp.bl == 0) {
      kst[2] = cs[0];
      kst[3] = cs[1];
    }
  }
  __syncthreads();
  const T a = (T)cs[0], a2 = (T)cs[1];
  for (int qq = q0 + (int)threadIdx.x; qq < q1; qq += kBlock) {
    const int P = qq * V;
    if (P != Pf) {
      Kc = ldv<T, V>(kc + P);
      D = ldv<T, V>(d + P);
    }
    VecN<T, V> O;
#pragma unroll
    for (int u = 0; u < V; u++) O.v[u] = a * Kc.v[u] + a2 * D.v[u];
    stv<T, V>(out + P, O);
  }
}

template <typename T, int V>
__global__ __launch_bounds__(kBlock) PSB_SMALL void k_kc_comb(Grid g, const T *kc, const T *d, T *out,
                                                    double *kst, int nprev,
                                                    const double *partial, const DevState *st) {
  if (skip_done(st, g.N)) return;
  kc_comb_body<T, V>(grid_span(), g, kc, d, out, kst, nprev, partial);
}

// ---------------------------------------------------------------------------
// the reference's additive-correction cycle (AC::v_cycle,
// additive_recursive.cpp:4-34): its restriction, interpolation and the
// re-initialisation of the Krylov state between the level solves
// ---------------------------------------------------------------------------

// AC::restriction (additive_restriction.cpp:12-177) for any coarsening
// factors: fH = sum_block f + external couplings * phi_outside - D phi, with
// D = c - internal couplings, in the reference's order of operations; 0 on
// inactive coarse cells.  Partials: sum fH and active count (for the mean).
static __global__ __launch_bounds__(kBlock) void k_ac_restrict(Grid g, DMat<double> A,
                                                        const double *phi, const double *f,
                                                        Grid gH, int CI, int CJ, int CK,
                                                        const uint8_t *afH, double *fH,
                                                        Work w) {
  double acc[2] = {0.0, 0.0};
  for (int Q = blockIdx.x * blockDim.x + threadIdx.x; Q < gH.N; Q += gridDim.x * blockDim.x) {
    int I, J, K;
    gH.ijk(Q, I, J, K);
    const int i0 = I * CI, j0 = J * CJ, k0 = K * CK;
    auto F = [&](int a, int b, int c) { return ((i0 + a) * g.ny + (j0 + b)) * g.nz + (k0 + c); };
    auto inside = [&](int a, int b, int c) {
      return i0 + a >= 0 && i0 + a < g.nx && j0 + b >= 0 && j0 + b < g.ny && k0 + c >= 0 &&
             k0 + c < g.nz;
    };
    auto PH = [&](int a, int b, int c) { return inside(a, b, c) ? phi[F(a, b, c)] : 0.0; };
    // e, n, t of a cell: its own (7 arrays) or, in symmetric storage, the
    // neighbour's w, s, b; 0 at walls
    auto e_ = [&](int a, int b, int c) {
      return i0 + a < g.nx - 1 ? (A.e ? A.e[F(a, b, c)] : A.w[F(a, b, c) + g.sI]) : 0.0;
    };
    auto n_ = [&](int a, int b, int c) {
      return j0 + b < g.ny - 1 ? (A.n ? A.n[F(a, b, c)] : A.s[F(a, b, c) + g.nz]) : 0.0;
    };
    auto t_ = [&](int a, int b, int c) {
      return k0 + c < g.nz - 1 ? (A.t ? A.t[F(a, b, c)] : A.b[F(a, b, c) + 1]) : 0.0;
    };
    double s = 0.0;
    for (int a = 0; a < CI; a++)
      for (int b = 0; b < CJ; b++)
        for (int c = 0; c < CK; c++) s += f[F(a, b, c)];
    for (int b = 0; b < CJ; b++)
      for (int c = 0; c < CK; c++) {
        s += e_(CI - 1, b, c) * PH(CI, b, c);
        s += A.w[F(0, b, c)] * PH(-1, b, c);
      }
    for (int a = 0; a < CI; a++)
      for (int c = 0; c < CK; c++) {
        s += n_(a, CJ - 1, c) * PH(a, CJ, c);
        s += A.s[F(a, 0, c)] * PH(a, -1, c);
      }
    for (int a = 0; a < CI; a++)
      for (int b = 0; b < CJ; b++) {
        s += t_(a, b, CK - 1) * PH(a, b, CK);
        s += A.b[F(a, b, 0)] * PH(a, b, -1);
      }
    for (int a = 0; a < CI; a++)
      for (int b = 0; b < CJ; b++)
        for (int c = 0; c < CK; c++) {
          // D of the cell: c minus its couplings inside the block, in the
          // reference's order (e, w, n, s, t, b)
          double d = A.c[F(a, b, c)];
          if (a < CI - 1) d -= e_(a, b, c);
          if (a > 0) d -= A.w[F(a, b, c)];
          if (b < CJ - 1) d -= n_(a, b, c);
          if (b > 0) d -= A.s[F(a, b, c)];
          if (c < CK - 1) d -= t_(a, b, c);
          if (c > 0) d -= A.b[F(a, b, c)];
          s -= d * phi[F(a, b, c)];
        }
    const bool act = !afH || afH[Q];
    fH[Q] = act ? s : 0.0;
    if (act) {
      acc[0] += s;
      acc[1] += 1.0;
    }
  }
  block_partial<2>(acc, w.partial);
}

// mean removal of the restricted right-hand side (additive_restriction.cpp:151-174)
static __global__ void k_ac_submean(Grid gH, double *fH, const uint8_t *afH, const DevState *st) {
  const double m = st->mean[0];
  for (int Q = blockIdx.x * blockDim.x + threadIdx.x; Q < gH.N; Q += gridDim.x * blockDim.x)
    if (!afH || afH[Q]) fH[Q] = fH[Q] - m;
}

// AC::interpolation (additive_interpolation.cpp:12-64): phi_h += phi_H
static __global__ void k_ac_interp(Grid g, double *phi, Grid gH, const double *phiH, int CI, int CJ,
                            int CK) {
  for (int P = blockIdx.x * blockDim.x + threadIdx.x; P < g.N; P += gridDim.x * blockDim.x) {
    int i, j, k;
    g.ijk(P, i, j, k);
    phi[P] = phi[P] + phiH[((i / CI) * gH.ny + (j / CJ)) * gH.nz + (k / CK)];
  }
}

// stopping parameters of the next level solve (Linear::solve arguments of
// AC::call_smoother / call_solver); the start kernel's finalize resets the
// Krylov scalars
static __global__ void k_stinit(DevState *st, int maxit, double res_rat, double res_tol, double scale,
                         double ntot, int stalecount) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  st->res_rat = res_rat;
  st->res_tol = res_tol;
  st->scale = scale;
  st->ntot = ntot;
  st->mini = 1;
  st->maxit = maxit;
  st->stalecount = stalecount;
  st->norm = 0;
  st->hist = nullptr;
  st->hist_cap = 0;
  st->done = 0;
}

// ---------------------------------------------------------------------------
// Pressure::discretize and Pressure::update_rhs on the device
// ---------------------------------------------------------------------------
struct DevMetrics {  // per direction: cell size, distance to the lower / upper centre
  const double *dxc, *dxw, *dxe, *dyc, *dys, *dyn, *dzc, *dzb, *dzt;
};

// pressure_discretize.cpp:24-148 with centered_system_bnd.cpp:27-61, one
// cell per thread: the six couplings dS / (d (rho_P + rho_nb)/2) from the
// density with its ghost layer, c their sum, couplings to solid cells
// zeroed on both sides (c of a fluid cell re-summed, of a solid cell kept),
// the boundary sides' couplings treated by their condition (bd: walls moved
// out of c, Dirichlet sides kept in c, their sources into F),
// c of solid cells scaled by 1e-12.  The same operations in the same order
// as the reference, so the matrix is bit-identical; stored symmetric
// (c, w, s, b: e, n, t are the neighbours' w, s, b).
static __global__ void k_discretize(Grid g, DevMetrics m, const double *rho, const uint8_t *off,
                                    psb_bnd bd, double *C, double *W, double *S, double *B,
                                    double *F) {
  const int gy = g.ny + 2, gz = g.nz + 2;
  for (int P = blockIdx.x * blockDim.x + threadIdx.x; P < g.N; P += gridDim.x * blockDim.x) {
    int i, j, k;
    g.ijk(P, i, j, k);
    auto R = [&](int a, int b, int c) {
      return rho[((size_t)(a + 1) * gy + (b + 1)) * gz + (c + 1)];
    };
    const double rp = R(i, j, k);
    const double a_x = m.dyc[j] * m.dzc[k];
    const double a_y = m.dxc[i] * m.dzc[k];
    const double a_z = m.dxc[i] * m.dyc[j];
    double w = a_x / m.dxw[i] / (0.5 * (rp + R(i - 1, j, k)));
    double e = a_x / m.dxe[i] / (0.5 * (R(i + 1, j, k) + rp));
    double s = a_y / m.dys[j] / (0.5 * (rp + R(i, j - 1, k)));
    double n = a_y / m.dyn[j] / (0.5 * (R(i, j + 1, k) + rp));
    double b = a_z / m.dzb[k] / (0.5 * (rp + R(i, j, k - 1)));
    double t = a_z / m.dzt[k] / (0.5 * (R(i, j, k + 1) + rp));
    double c = w + e + s + n + b + t;
    if (off) {
      if (off[P]) {
        w = e = s = n = b = t = 0.0;
      } else {
        if (i > 0 && off[P - g.sI]) w = 0.0;
        if (i < g.nx - 1 && off[P + g.sI]) e = 0.0;
        if (j > 0 && off[P - g.nz]) s = 0.0;
        if (j < g.ny - 1 && off[P + g.nz]) n = 0.0;
        if (k > 0 && off[P - 1]) b = 0.0;
        if (k < g.nz - 1 && off[P + 1]) t = 0.0;
        c = w + e + s + n + b + t;
      }
    }
    // boundary sides (centered_system_bnd.cpp:5-61, 119-160): the Neumann
    // group (wall, symmetry, outlet, neumann) in the order imin .. kmax moves
    // the side's coupling out of c and adds value * dS to fbnd; then the
    // Dirichlet group (dirichlet, inlet) keeps the coupling in c and adds
    // value * coupling
    const bool at[6] = {i == 0, i == g.nx - 1, j == 0, j == g.ny - 1, k == 0, k == g.nz - 1};
    double cp[6] = {w, e, s, n, b, t};
    const double dS[6] = {a_x, a_x, a_y, a_y, a_z, a_z};
    double fb = 0.0;
    for (int grp = 0; grp < 2; grp++)
      for (int q = 0; q < 6; q++) {
        if (!at[q] || bd.type[q] != grp) continue;
        if (grp == 0) {
          fb += bd.value[q] * dS[q];
          c -= cp[q];
        } else {
          fb += bd.value[q] * cp[q];
        }
        cp[q] = 0.0;
      }
    w = cp[0];
    s = cp[2];
    b = cp[4];
    if (off && off[P]) c *= 1.0e-12;  // Src/Global/global_constants.cpp:17 (pico)
    C[P] = c;
    W[P] = w;
    S[P] = s;
    B[P] = b;
    if (F) F[P] = fb;
  }
}

// pressure_update_rhs.cpp:27-52 with dt = 1: b = (u_w - u_e) dSx + (v_s -
// v_n) dSy + (w_b - w_t) dSz from the staggered face velocities (u of
// (nx+1) x ny x nz faces, v of nx x (ny+1) x nz, w of nx x ny x (nz+1)); 0 on
// solid cells; plus the boundary source fbnd when the assembly formed one.  Partials: sum b and fluid cells (for the optional zero mean).
static __global__ __launch_bounds__(kBlock) void k_update_rhs(Grid g, DevMetrics m,
                                                              const double *u, const double *v,
                                                              const double *wv,
                                                              const uint8_t *off,
                                                              const double *fbnd, double *bout,
                                                              Work w) {
  double acc[2] = {0.0, 0.0};
  for (int P = blockIdx.x * blockDim.x + threadIdx.x; P < g.N; P += gridDim.x * blockDim.x) {
    int i, j, k;
    g.ijk(P, i, j, k);
    const double a_x = m.dyc[j] * m.dzc[k] * 1.0;
    const double a_y = m.dxc[i] * m.dzc[k] * 1.0;
    const double a_z = m.dxc[i] * m.dyc[j] * 1.0;
    const size_t ny = g.ny, nz = g.nz;
    double f = 0.0;
    f += a_x * u[((size_t)i * ny + j) * nz + k];
    f -= a_x * u[((size_t)(i + 1) * ny + j) * nz + k];
    f += a_y * v[((size_t)i * (ny + 1) + j) * nz + k];
    f -= a_y * v[((size_t)i * (ny + 1) + j + 1) * nz + k];
    f += a_z * wv[((size_t)i * ny + j) * (nz + 1) + k];
    f -= a_z * wv[((size_t)i * ny + j) * (nz + 1) + k + 1];
    const bool solid = off && off[P];
    if (solid) f = 0.0;
    if (fbnd) f += 0.0 + fbnd[P];  // fnew += fext + fbnd (:135), fext 0 here
    bout[P] = f;
    if (!solid) {
      acc[0] += f;
      acc[1] += 1.0;
    }
  }
  block_partial<2>(acc, w.partial);
}

// b -= mean over the fluid cells (the benchmark's zero-mean right-hand side)
static __global__ void k_rhs_center(Grid g, double *b, const uint8_t *off, const DevState *st) {
  const double mean = st->mean[0];
  for (int P = blockIdx.x * blockDim.x + threadIdx.x; P < g.N; P += gridDim.x * blockDim.x)
    if (!(off && off[P])) b[P] = b[P] - mean;
}

// Momentum::project (momentum_project.cpp:10-76), one cell per thread: the
// cell's west, south and bottom faces when they are interior faces (the wall
// faces stay, staggered_set_indices.cpp:9-19) take
// (p_nb - p_P) / rho_f * dt / d with rho_f = (rho_P + rho_nb) / 2
// (property_value.cpp:18-32) and d the centre distance (dxw, dys, dzb).
// With solid cells (off) the body form: faces between two fluid cells only,
// rounded as (p_nb - p_P) * dt / rho_f / d (:50-70).  mask bits 0/1/2 select
// u, v, w (Momentum's ifull / jfull / kfull).  64 B per cell: p, rho, and
// the three face arrays read and written.
static __global__ __launch_bounds__(kBlock) void k_project(Grid g, DevMetrics m,
                                                           const double *rho, const uint8_t *off,
                                                           const double *p, double dt, int mask,
                                                           double *u, double *v, double *wv) {
  const int gy = g.ny + 2, gz = g.nz + 2;
  const size_t ny = g.ny, nz = g.nz;
  for (int P = blockIdx.x * blockDim.x + threadIdx.x; P < g.N; P += gridDim.x * blockDim.x) {
    int i, j, k;
    g.ijk(P, i, j, k);
    auto R = [&](int a, int b, int c) {
      return rho[((size_t)(a + 1) * gy + (b + 1)) * gz + (c + 1)];
    };
    const double pp = p[P], rp = R(i, j, k);
    const bool fluid = !(off && off[P]);
    if ((mask & 1) && i > 0) {
      const int Q = P - g.sI;
      const double rf = 0.5 * (rp + R(i - 1, j, k));
      double &f = u[P];  // face i of (nx+1) x ny x nz: same index as the cell
      if (!off) f += (p[Q] - pp) / rf * dt / m.dxw[i];
      else if (fluid && !off[Q]) f += (p[Q] - pp) * dt / rf / m.dxw[i];
    }
    if ((mask & 2) && j > 0) {
      const int Q = P - g.nz;
      const double rf = 0.5 * (rp + R(i, j - 1, k));
      double &f = v[((size_t)i * (ny + 1) + j) * nz + k];
      if (!off) f += (p[Q] - pp) / rf * dt / m.dys[j];
      else if (fluid && !off[Q]) f += (p[Q] - pp) * dt / rf / m.dys[j];
    }
    if ((mask & 4) && k > 0) {
      const int Q = P - 1;
      const double rf = 0.5 * (rp + R(i, j, k - 1));
      double &f = wv[((size_t)i * ny + j) * (nz + 1) + k];
      if (!off) f += (p[Q] - pp) / rf * dt / m.dzb[k];
      else if (fluid && !off[Q]) f += (p[Q] - pp) * dt / rf / m.dzb[k];
    }
  }
}

// Pressure::ghost (pressure_ghost.cpp:10-190), one cell per thread: each of
// the six faces (w, e, s, n, b, t, in that order) that the colour function
// crosses at 0.5 adds iphase dS sigma kappa_f / (rho_f d) to fext, iphase =
// -1 below 0.5; rho_f = (rho_hi + rho_lo) / 2 of the face's upper and lower
// cell (property_value.cpp:59-83), kappa_f their harmonic mean when both
// have one sign, else the arithmetic one (:192-213); d = dxw, dxe, dys, dys
// (the reference's north face uses dys, :74), dzb, dzt.  c, kappa, rho with
// their ghost layer; solid cells are skipped.  The reference's operations in
// its order (bit-identical with -ffp-contract=off).
static __global__ __launch_bounds__(kBlock) void k_ghost(Grid g, DevMetrics m, const double *c,
                                                         const double *kap, const double *rho,
                                                         double sigma, const uint8_t *off,
                                                         double *fext) {
  const int gy = g.ny + 2, gz = g.nz + 2;
  for (int P = blockIdx.x * blockDim.x + threadIdx.x; P < g.N; P += gridDim.x * blockDim.x) {
    if (off && off[P]) continue;
    int i, j, k;
    g.ijk(P, i, j, k);
    auto at = [&](const double *a, int di, int dj, int dk) {
      return a[((size_t)(i + di + 1) * gy + (j + dj + 1)) * gz + (k + dk + 1)];
    };
    const double cint = 0.5, cc = at(c, 0, 0, 0);
    const double iphase = cc < cint ? -1.0 : 1.0;
    double f = fext[P];
    auto face = [&](int di, int dj, int dk, double area, double d) {
      const double cn = at(c, di, dj, dk);
      if (!((cc - cint) * (cn - cint) < 0.0)) return;
      const bool lo = di + dj + dk < 0;  // the neighbour is the face's lower cell
      const double rn = at(rho, di, dj, dk), rp = at(rho, 0, 0, 0);
      const double kn = at(kap, di, dj, dk), kp = at(kap, 0, 0, 0);
      const double r0 = lo ? rn : rp, r1 = lo ? rp : rn;
      const double k0 = lo ? kn : kp, k1 = lo ? kp : kn;
      const double rf = 0.5 * (r1 + r0);
      double kf = 0.5 * (k0 + k1);
      if (k0 * k1 > 0.0) kf = 2.0 * k0 * k1 / (k0 + k1);
      f += iphase * area * sigma * kf / (rf * d);
    };
    const double ax = m.dyc[j] * m.dzc[k], ay = m.dxc[i] * m.dzc[k], az = m.dxc[i] * m.dyc[j];
    face(-1, 0, 0, ax, m.dxw[i]);
    face(1, 0, 0, ax, m.dxe[i]);
    face(0, -1, 0, ay, m.dys[j]);
    face(0, 1, 0, ay, m.dys[j]);
    face(0, 0, -1, az, m.dzb[k]);
    face(0, 0, 1, az, m.dzt[k]);
    fext[P] = f;
  }
}

// ---------------------------------------------------------------------------
// set-up kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T *x, T v, int n) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) x[q] = v;
}

// plain copy of n floats (n % 4 == 0 for the vector part): the reference
// stream for calibrating hardware byte counters
static __global__ __launch_bounds__(kBlock) void k_copy(const float *x, float *y, int n) {
  const int nv = n / 4;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nv; q += gridDim.x * blockDim.x)
    stv<float, 4>(y + 4 * q, ldv<float, 4>(x + 4 * q));
  for (int q = 4 * nv + blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
    y[q] = x[q];
}

// b = (half)(scale * a): the smoother's copy of a level matrix.  A positive
// diagonal (pos = 1) never rounds to zero: the solid cells' c (scaled by
// 1e-12) is raised to the smallest normal half, 2^-14.
static __global__ void k_to_f16(const double *a, _Float16 *b, double scale, int n, int pos) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    float v = (float)(scale * a[q]);
    if (pos && a[q] > 0.0 && v < 6.103515625e-05f) v = 6.103515625e-05f;
    b[q] = (_Float16)v;
  }
}

// d = c - (w + e + s + n + b + t) of a symmetric 4-array FP64 level matrix
// (e, n, t read from the neighbours), rounded to T: the diagonal of the
// flux form
template <typename T>
__global__ void k_flux_d(Grid g, DMat<double> A, T *d, double scale) {
  for (int P = blockIdx.x * blockDim.x + threadIdx.x; P < g.N; P += gridDim.x * blockDim.x) {
    int i, j, k;
    g.ijk(P, i, j, k);
    double s = A.w[P] + (i < g.nx - 1 ? A.w[P + g.sI] : 0.0);
    s += A.s[P] + (j < g.ny - 1 ? A.s[P + g.nz] : 0.0);
    s += A.b[P] + (k < g.nz - 1 ? A.b[P + 1] : 0.0);
    d[P] = (T)(float)(scale * (A.c[P] - s));
  }
}

// per-block max |r / dV| (Scalar::max_abs_voldiv, scalar_minmax.cpp:155-165;
// the maximum does not depend on the order) into out[blockIdx.x]
static __global__ __launch_bounds__(kBlock) void k_absmax_voldiv(Grid g, Metric m,
                                                                 const double *r, double *out) {
  __shared__ double sm[kBlock];
  double v = 0.0;
  for (int P = blockIdx.x * blockDim.x + threadIdx.x; P < g.N; P += gridDim.x * blockDim.x) {
    int i, j, k;
    g.ijk(P, i, j, k);
    v = fmax(v, fabs(r[P] / m.dV(i, j, k)));
  }
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sm[0];
}

// one block: out[0] = max of a[0..n)
static __global__ __launch_bounds__(kBlock) void k_max_final(const double *a, int n,
                                                             double *out) {
  __shared__ double sm[kBlock];
  double v = 0.0;
  for (int q = threadIdx.x; q < n; q += blockDim.x) v = fmax(v, a[q]);
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sm[0];
}

// per-block maximum of a (n >= 1) into out[blockIdx.x]
static __global__ __launch_bounds__(kBlock) void k_blockmax(const double *a, int n,
                                                            double *out) {
  __shared__ double sm[kBlock];
  double m = -1.0e308;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
    m = fmax(m, a[q]);
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sm[0];
}

static __global__ void k_to_f32(const double *a, float *b, int n) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
    b[q] = (float)a[q];
}

// Galerkin coarse matrix for piecewise-constant prolongation, one thread per
// coarse cell; the fine block is visited in the order of
// additive_coarsen_system.cpp:103-219 (c summed, internal e/w per (j,k),
// n/s per (i,k), t/b per (i,j); external couplings summed over the faces).
// The fine matrix is symmetric 4-array; so is the coarse one.
static __global__ void k_galerkin(Grid g, DMat<double> A, Grid gH, int CI, int CJ, int CK,
                           double *cH, double *wH, double *sH, double *bH, double *eH,
                           double *nH, double *tH) {
  for (int Q = blockIdx.x * blockDim.x + threadIdx.x; Q < gH.N; Q += gridDim.x * blockDim.x) {
    int I, J, K;
    gH.ijk(Q, I, J, K);
    auto F = [&](int a, int b, int c) {
      return ((I * CI + a) * g.ny + (J * CJ + b)) * g.nz + (K * CK + c);
    };
    // e, n, t of a fine cell: its own (7 arrays) or the neighbour's w, s, b
    auto e_ = [&](int a, int b, int c) {
      return I * CI + a < g.nx - 1 ? (A.e ? A.e[F(a, b, c)] : A.w[F(a, b, c) + g.sI]) : 0.0;
    };
    auto n_ = [&](int a, int b, int c) {
      return J * CJ + b < g.ny - 1 ? (A.n ? A.n[F(a, b, c)] : A.s[F(a, b, c) + g.nz]) : 0.0;
    };
    auto t_ = [&](int a, int b, int c) {
      return K * CK + c < g.nz - 1 ? (A.t ? A.t[F(a, b, c)] : A.b[F(a, b, c) + 1]) : 0.0;
    };
    double c = 0.0;
    for (int a = 0; a < CI; a++)
      for (int b = 0; b < CJ; b++)
        for (int cc = 0; cc < CK; cc++) c += A.c[F(a, b, cc)];
    for (int b = 0; b < CJ; b++)
      for (int cc = 0; cc < CK; cc++) {
        for (int a = 0; a < CI - 1; a++) c -= e_(a, b, cc);
        for (int a = 1; a < CI; a++) c -= A.w[F(a, b, cc)];
      }
    double w = 0.0, e = 0.0;
    for (int b = 0; b < CJ; b++)
      for (int cc = 0; cc < CK; cc++) {
        e += e_(CI - 1, b, cc);
        w += A.w[F(0, b, cc)];
      }
    for (int a = 0; a < CI; a++)
      for (int cc = 0; cc < CK; cc++) {
        for (int b = 0; b < CJ - 1; b++) c -= n_(a, b, cc);
        for (int b = 1; b < CJ; b++) c -= A.s[F(a, b, cc)];
      }
    double s = 0.0, n = 0.0;
    for (int a = 0; a < CI; a++)
      for (int cc = 0; cc < CK; cc++) {
        n += n_(a, CJ - 1, cc);
        s += A.s[F(a, 0, cc)];
      }
    for (int a = 0; a < CI; a++)
      for (int b = 0; b < CJ; b++) {
        for (int cc = 0; cc < CK - 1; cc++) c -= t_(a, b, cc);
        for (int cc = 1; cc < CK; cc++) c -= A.b[F(a, b, cc)];
      }
    double bb = 0.0, t = 0.0;
    for (int a = 0; a < CI; a++)
      for (int b = 0; b < CJ; b++) {
        t += t_(a, b, CK - 1);
        bb += A.b[F(a, b, 0)];
      }
    cH[Q] = c;
    wH[Q] = w;
    sH[Q] = s;
    bH[Q] = bb;
    if (eH) {  // non-symmetric levels keep their own e, n, t
      eH[Q] = e;
      nH[Q] = n;
      tH[Q] = t;
    }
  }
}

// active flag of a coarse cell: OR over its block (additive_coarsen_flag.cpp)
static __global__ void k_coarsen_flag(Grid g, const uint8_t *f, Grid gH, int CI, int CJ, int CK,
                               uint8_t *fH) {
  for (int Q = blockIdx.x * blockDim.x + threadIdx.x; Q < gH.N; Q += gridDim.x * blockDim.x) {
    int I, J, K;
    gH.ijk(Q, I, J, K);
    int fl = 0;
    for (int a = 0; a < CI; a++)
      for (int b = 0; b < CJ; b++)
        for (int cc = 0; cc < CK; cc++)
          fl |= f[((I * CI + a) * g.ny + (J * CJ + b)) * g.nz + (K * CK + cc)];
    fH[Q] = (uint8_t)(fl != 0);
  }
}

}  // namespace psb
