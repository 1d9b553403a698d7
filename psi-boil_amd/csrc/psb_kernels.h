// psb_kernels.h -- the hand-written gfx950 kernels of the pressure solver.
//
// Every kernel streams the interior arrays once (HBM-bound; no MFMA): a
// block of 256 threads walks a contiguous run of V-wide cell vectors
// (V = 16 bytes / sizeof(T)), neighbours come through L1/L2, and the
// XCD-aware block remap keeps a slab's i+-1 planes in one L2.  Reductions
// are deterministic (fixed-order last-block sums) and accumulate in FP64 for
// both precisions.  Kernels of an iteration start with `if (st->done)
// return;`, so a captured graph of several iterations stops doing work as
// soon as the solver has finished.
#pragma once

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
  double *partial;
  unsigned *counter;
};

#define PSB_TILE_BEGIN(g, V)                                                  \
  const int nvec_ = (g).N / (V);                                              \
  const int bl_ = xcd_block();                                                \
  const int per_ = (nvec_ + (int)gridDim.x - 1) / (int)gridDim.x;             \
  const int q0_ = bl_ * per_;                                                 \
  const int q1_ = min(q0_ + per_, nvec_);                                     \
  for (int q_ = q0_ + (int)threadIdx.x; q_ < q1_; q_ += kBlock) {             \
    const int P = q_ * (V);                                                   \
    const int i = P / (g).sI;                                                 \
    const int rem_ = P - i * (g).sI;                                          \
    const int j = rem_ / (g).nz;                                              \
    const int k = rem_ - j * (g).nz;                                          \
    (void)i; (void)j; (void)k;
#define PSB_TILE_END }

// stop logic shared by the Krylov loops (cg.cpp:121-160, bicgs.cpp:174-210):
// called by one thread after iteration st->iter with the new residual.
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
// Krylov start: r = b - A x, res0, rho0 and (AC) e1 = omega r / c.
// mode 0: diagonal CG (rho = r.(r/c)), 1: preconditioned CG (e1),
// 2: BiCGS (r~ = r, rho = r~.r).
// ---------------------------------------------------------------------------
template <typename T, int V, bool SYM, int MODE>
__global__ __launch_bounds__(kBlock) void k_start(Grid g, DMat<T> A, Metric m, const T *b,
                                                  const T *x, T *r, T *aux, T omega,
                                                  DevState *st, Work w) {
  ArrX<T, V> X{x};
  double acc[3] = {0.0, 0.0, 0.0};  // sum r^2/dV^2, sum r^2, rho
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
  double tot[3];
  if (grid_sum<3>(acc, w.partial, w.counter, tot) && threadIdx.x == 0) {
    double res = st->norm == 0 ? ref_norm(st, tot[0]) : sqrt(tot[1]);
    st->res0 = res;
    st->res = res;
    st->rsq0 = tot[1];
    st->rsq = tot[1];
    st->rho = tot[2];
    st->beta = 0.0;
    st->iter = 0;
    st->report_iters = 0;
    st->converged = 0;
    st->reason = 0;
    st->done = 0;
    if (st->hist && st->hist_cap > 0) st->hist[0] = res;
    for (int q = 0; q < st->stalecount && q < kStaleMax; q++) st->stale[q] = 1.0e24;
    if (st->stalecount > 0) st->stale[st->stalecount - 1] = res;
    if (res < st->res_tol) {  // cg.cpp:64 early return
      st->done = 1;
      st->converged = 1;
      st->reason = 7;
    }
    if (!(res == res)) {
      st->done = 1;
      st->reason = 8;
    }
    if (MODE == 2 && !st->done && sqrt(fabs(tot[2])) < st->res_tol) {
      st->done = 1;  // bicgs.cpp:89 at i = 0
      st->reason = 4;
      st->report_iters = 1;
    }
  }
}

// ---------------------------------------------------------------------------
// CG step 1: p = z + beta p_old (z = r/c for the diagonal preconditioner,
// else the preconditioner output), q = A p, alpha = rho / (p.q).
// cg.cpp:75-108
// ---------------------------------------------------------------------------
template <typename T, int V, bool SYM, bool DIAG>
__global__ __launch_bounds__(kBlock) void k_cg_pq(Grid g, DMat<T> A, const T *r_or_z,
                                                  const T *pold, T *pnew, T *q,
                                                  DevState *st, Work w) {
  if (st->done) return;
  const T beta = (T)st->beta;
  double acc[1] = {0.0};
  PSB_TILE_BEGIN(g, V)
  Gath<T, V> G;
  if (DIAG) {
    DiagPX<T, V> X{r_or_z, A.c, pold, beta};
    G = gather<T, V, SYM>(g, A, X, P, i, j, k);
  } else {
    ZPX<T, V> X{r_or_z, pold, beta};
    G = gather<T, V, SYM>(g, A, X, P, i, j, k);
  }
  VecN<T, V> Q = apply_A<T, V>(G);
  stv<T, V>(pnew + P, G.x);
  stv<T, V>(q + P, Q);
#pragma unroll
  for (int u = 0; u < V; u++) acc[0] += (double)(G.x.v[u] * Q.v[u]);
  PSB_TILE_END
  double tot[1];
  if (grid_sum<1>(acc, w.partial, w.counter, tot) && threadIdx.x == 0) {
    st->pq = tot[0];
    st->alpha = st->rho / tot[0];
  }
}

// ---------------------------------------------------------------------------
// CG step 2: x += alpha p, r -= alpha q, residual norm and stop test
// (cg.cpp:114-160); diagonal mode also forms rho_next = r.(r/c) and beta;
// preconditioned mode writes e1 = omega r / c, the first Jacobi sweep of the
// next V-cycle.
// ---------------------------------------------------------------------------
template <typename T, int V, bool DIAG>
__global__ __launch_bounds__(kBlock) void k_cg_xr(Grid g, const T *c, Metric m, const T *p,
                                                  const T *q, T *x, T *r, T *e1, T omega,
                                                  DevState *st, Work w) {
  if (st->done) return;
  const T alpha = (T)st->alpha;
  double acc[3] = {0.0, 0.0, 0.0};  // voldiv, r^2, rho_next
  PSB_TILE_BEGIN(g, V)
  VecN<T, V> X = ldv<T, V>(x + P), R = ldv<T, V>(r + P), Pp = ldv<T, V>(p + P),
             Q = ldv<T, V>(q + P), C = ldv<T, V>(c + P), E;
#pragma unroll
  for (int u = 0; u < V; u++) {
    X.v[u] = X.v[u] + alpha * Pp.v[u];
    R.v[u] = R.v[u] - alpha * Q.v[u];
    double rd = (double)R.v[u];
    double dv = m.dV(i, j, k + u);
    acc[0] += rd * rd / dv / dv;
    acc[1] += rd * rd;
    if (DIAG) {
      T z = R.v[u] * (T(1) / C.v[u]);
      acc[2] += (double)(R.v[u] * z);
    } else {
      E.v[u] = omega * R.v[u] / C.v[u];
    }
  }
  stv<T, V>(x + P, X);
  stv<T, V>(r + P, R);
  if (!DIAG) stv<T, V>(e1 + P, E);
  PSB_TILE_END
  double tot[3];
  if (grid_sum<3>(acc, w.partial, w.counter, tot) && threadIdx.x == 0) {
    double res = st->norm == 0 ? ref_norm(st, tot[0]) : sqrt(tot[1]);
    st->rsq = tot[1];
    krylov_stop(st, res, false);
    if (DIAG && !st->done) {
      st->beta = tot[2] / st->rho;  // cg.cpp:93
      st->rho = tot[2];
    }
  }
}

// ---------------------------------------------------------------------------
// damped Jacobi sweep  e_out = e + omega (r - A e) / c, with the operand
// either e itself or e + cfac P eH (prolongation fused into the first post
// sweep).  RHO: also form rho = r.z for the CG step that follows (last post
// sweep on the fine level).
// ---------------------------------------------------------------------------
template <typename T, int V, bool SYM, bool PROL, bool RHO>
__global__ __launch_bounds__(kBlock) void k_jacobi(Grid g, DMat<T> A, const T *r, const T *e,
                                                   T *eout, ProlX<T, V> pr, T omega,
                                                   DevState *st, Work w) {
  if (st->done) return;
  double acc[1] = {0.0};
  PSB_TILE_BEGIN(g, V)
  Gath<T, V> G;
  if (PROL) {
    G = gather<T, V, SYM>(g, A, pr, P, i, j, k);
  } else {
    ArrX<T, V> X{e};
    G = gather<T, V, SYM>(g, A, X, P, i, j, k);
  }
  VecN<T, V> R = ldv<T, V>(r + P);
  VecN<T, V> t = resid<T, V>(R, G), E;
#pragma unroll
  for (int u = 0; u < V; u++) {
    E.v[u] = G.x.v[u] + omega * t.v[u] / G.c.v[u];
    if (RHO) acc[0] += (double)(R.v[u] * E.v[u]);
  }
  stv<T, V>(eout + P, E);
  PSB_TILE_END
  if (RHO) {
    double tot[1];
    if (grid_sum<1>(acc, w.partial, w.counter, tot) && threadIdx.x == 0) {
      st->beta = st->iter == 0 ? 0.0 : tot[0] / st->rho;
      st->rho = tot[0];
    }
  }
}

// ---------------------------------------------------------------------------
// restriction for the preconditioner: rH[I] = sum over the fine block of
// (r - A e)  (the residual sum of additive_restriction.cpp:12-135 without
// the reference's re-expansion; solid fine cells left out, the transpose of
// the masked prolongation), 0 on inactive coarse cells, and the sum
// over active cells for the mean removal of additive_restriction.cpp:151-174.
// One thread per coarse cell.
// ---------------------------------------------------------------------------
template <typename T, bool SYM>
__global__ __launch_bounds__(kBlock) void k_restrict(Grid g, DMat<T> A, const T *r, const T *e,
                                                     const uint8_t *aflag, Grid gH, int CI,
                                                     int CJ, int CK, const uint8_t *aflagH,
                                                     T *rH, int lev, DevState *st, Work w) {
  if (st->done) return;
  ArrX<T, 1> X{e};
  double acc[2] = {0.0, 0.0};
  const int b = xcd_block();
  for (int Q = b * kBlock + threadIdx.x; Q < gH.N; Q += gridDim.x * kBlock) {
    const int I = Q / gH.sI, remQ = Q - I * gH.sI, J = remQ / gH.nz, K = remQ - J * gH.nz;
    T s = T(0);
    for (int a = 0; a < CI; a++)
      for (int bb = 0; bb < CJ; bb++)
        for (int cc = 0; cc < CK; cc++) {
          const int i = I * CI + a, j = J * CJ + bb, k = K * CK + cc;
          const int P = (i * g.ny + j) * g.nz + k;
          if (aflag && !aflag[P]) continue;
          Gath<T, 1> G = gather<T, 1, SYM>(g, A, X, P, i, j, k);
          s += resid<T, 1>(ldv<T, 1>(r + P), G).v[0];
        }
    const bool act = aflagH ? aflagH[Q] != 0 : true;
    if (!act) s = T(0);
    rH[Q] = s;
    if (act) {
      acc[0] += (double)s;
      acc[1] += 1.0;
    }
  }
  double tot[2];
  if (grid_sum<2>(acc, w.partial, w.counter, tot) && threadIdx.x == 0) {
    st->msum[lev] = tot[0];
    st->mean[lev] = tot[1] > 0.0 ? tot[0] / tot[1] : 0.0;
  }
}

// first sweep on a coarse level: rH -= mean (active cells), e1 = omega rH / c
template <typename T, int V>
__global__ __launch_bounds__(kBlock) void k_coarse_first(Grid g, const T *c, const uint8_t *aflag,
                                                         T *r, T *e1, T omega, int lev,
                                                         int nu1, DevState *st) {
  if (st->done) return;
  const T mean = (T)st->mean[lev];
  PSB_TILE_BEGIN(g, V)
  VecN<T, V> R = ldv<T, V>(r + P), C = ldv<T, V>(c + P), E;
#pragma unroll
  for (int u = 0; u < V; u++) {
    if (!aflag || aflag[P + u]) R.v[u] = R.v[u] - mean;
    E.v[u] = nu1 > 0 ? omega * R.v[u] / C.v[u] : T(0);
  }
  stv<T, V>(r + P, R);
  stv<T, V>(e1 + P, E);
  PSB_TILE_END
}

// coarsest level: z = G r, G the dense projected pseudo-inverse (FP64).
// One wave per row; r is read in T and promoted.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gemv(int m, const double *Gm, const T *r, T *z,
                                                 DevState *st) {
  if (st->done) return;
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (row >= m) return;
  const double *g = Gm + (size_t)row * m;
  double s = 0.0;
  for (int c = lane; c < m; c += 64) s += g[c] * (double)r[c];
  s = wave_sum(s);
  if (lane == 0) z[row] = (T)s;
}

// ---------------------------------------------------------------------------
// iterative refinement: r32 = (float)(b - A x) in FP64 with its FP64 norm,
// the inner solve's first Jacobi sweep e1 = omega r / c, and the FP32 norm
// that starts the inner solve.
// ---------------------------------------------------------------------------
template <int V, bool SYM>
__global__ __launch_bounds__(kBlock) void k_ir_resid(Grid g, DMat<double> A, Metric m,
                                                     const double *b, const double *x,
                                                     float *r32, float *e1, float omega,
                                                     DevState *st64, DevState *st32, Work w) {
  ArrX<double, V> X{x};
  double acc[3] = {0.0, 0.0, 0.0};  // r64^2, r32^2, r64^2/dV^2
  PSB_TILE_BEGIN(g, V)
  Gath<double, V> G = gather<double, V, SYM>(g, A, X, P, i, j, k);
  VecN<double, V> R = resid<double, V>(ldv<double, V>(b + P), G);
#pragma unroll
  for (int u = 0; u < V; u++) {
    float rf = (float)R.v[u];
    float cf = (float)G.c.v[u];
    r32[P + u] = rf;
    e1[P + u] = omega * rf / cf;
    double dv = m.dV(i, j, k + u);
    acc[0] += R.v[u] * R.v[u];
    acc[1] += (double)rf * (double)rf;
    acc[2] += R.v[u] * R.v[u] / dv / dv;
  }
  PSB_TILE_END
  double tot[3];
  if (grid_sum<3>(acc, w.partial, w.counter, tot) && threadIdx.x == 0) {
    st64->rsq = tot[0];
    st64->res = st64->norm == 0 ? ref_norm(st64, tot[2]) : sqrt(tot[0]);
    double r0 = sqrt(tot[1]);
    st32->res0 = r0;
    st32->res = r0;
    st32->rsq0 = tot[1];
    st32->rsq = tot[1];
    st32->beta = 0.0;
    st32->iter = 0;
    st32->report_iters = 0;
    st32->converged = 0;
    st32->reason = 0;
    st32->done = (r0 == 0.0) ? 1 : 0;
    if (st32->hist && st32->hist_cap > 0) st32->hist[0] = r0;
  }
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

// ---------------------------------------------------------------------------
// BiCGS (bicgs.cpp:79-216) in three passes per iteration.
// B1: p = r + beta (p_old - omega v_old) formed as the reference does
//     (p = r + beta p; p -= (beta omega) v), p^ = p / c (Diagonal), v = A p^,
//     alpha = rho / (r~.v).
// ---------------------------------------------------------------------------
template <typename T, int V>
struct BiPX {
  const T *r, *p, *v, *c;
  T beta, bo;
  __device__ __forceinline__ T pnew(T rr, T pp, T vv) const {
    T t = rr + beta * pp;
    return t - bo * vv;
  }
  __device__ __forceinline__ T one(int Q) const {
    return pnew(r[Q], p[Q], v[Q]) * (T(1) / c[Q]);
  }
  __device__ __forceinline__ VecN<T, V> vec(int P) const {
    VecN<T, V> R = ldv<T, V>(r + P), Pp = ldv<T, V>(p + P), Vv = ldv<T, V>(v + P),
               C = ldv<T, V>(c + P), o;
#pragma unroll
    for (int u = 0; u < V; u++) o.v[u] = pnew(R.v[u], Pp.v[u], Vv.v[u]) * (T(1) / C.v[u]);
    return o;
  }
  __device__ __forceinline__ T at(int Q) const { return one(Q); }
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
  double tot[1];
  if (grid_sum<1>(acc, w.partial, w.counter, tot) && threadIdx.x == 0) {
    st->rv = tot[0];
    st->alpha = st->rho / tot[0];
  }
}

// B2: s = r - alpha v, s^ = s / c, t = A s^; sums s.s, t.s, t.t; omega.
template <typename T, int V>
struct BiSX {
  const T *r, *v, *c;
  T alpha;
  __device__ __forceinline__ T s1(T rr, T vv) const { return rr - alpha * vv; }
  __device__ __forceinline__ VecN<T, V> vec(int P) const {
    VecN<T, V> R = ldv<T, V>(r + P), Vv = ldv<T, V>(v + P), C = ldv<T, V>(c + P), o;
#pragma unroll
    for (int u = 0; u < V; u++) o.v[u] = s1(R.v[u], Vv.v[u]) * (T(1) / C.v[u]);
    return o;
  }
  __device__ __forceinline__ T at(int Q) const { return s1(r[Q], v[Q]) * (T(1) / c[Q]); }
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
  double tot[3];
  if (grid_sum<3>(acc, w.partial, w.counter, tot) && threadIdx.x == 0) {
    st->rv_ss = tot[0];
    st->rv_ts = tot[1];
    st->rv_tt = tot[2];
    st->omega = tot[1] / tot[2];
  }
}

// B3: x += alpha p^; x += omega s^; r = s - omega t; res; rho_next = r~.r;
// stop tests in the reference's order, including the small-s exit
// (bicgs.cpp:134-138) where only x += alpha p^ is applied.
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
  double tot[3];
  if (grid_sum<3>(acc, w.partial, w.counter, tot) && threadIdx.x == 0) {
    if (small_s) {
      st->done = 1;
      st->converged = 0;
      st->reason = 5;
      st->report_iters = st->iter + 1;
      return;
    }
    double res = st->norm == 0 ? ref_norm(st, tot[0]) : sqrt(tot[1]);
    st->rsq = tot[1];
    const bool small_om = st->omega >= -st->res_tol && st->omega <= st->res_tol;
    krylov_stop(st, res, small_om);
    if (!st->done) {
      // top of the next iteration: rho = r~.r, breakdown test (bicgs.cpp:84-89)
      double rho = tot[2];
      if (sqrt(fabs(rho)) < st->res_tol) {
        st->done = 1;
        st->reason = 4;
        st->report_iters = st->iter + 1;
      } else {
        st->beta = (rho / st->rho) * (st->alpha / st->omega);  // bicgs.cpp:102
        st->rho = rho;
      }
    }
  }
}

// true relative residual ||b - A x|| / ||b|| (FP64 diagnostic after a solve)
template <int V, bool SYM>
__global__ __launch_bounds__(kBlock) void k_relres(Grid g, DMat<double> A, const double *b,
                                                   const double *x, double *out, Work w) {
  ArrX<double, V> X{x};
  double acc[2] = {0.0, 0.0};
  PSB_TILE_BEGIN(g, V)
  Gath<double, V> G = gather<double, V, SYM>(g, A, X, P, i, j, k);
  VecN<double, V> B = ldv<double, V>(b + P), R = resid<double, V>(B, G);
#pragma unroll
  for (int u = 0; u < V; u++) {
    acc[0] += R.v[u] * R.v[u];
    acc[1] += B.v[u] * B.v[u];
  }
  PSB_TILE_END
  double tot[2];
  if (grid_sum<2>(acc, w.partial, w.counter, tot) && threadIdx.x == 0) {
    out[0] = tot[0];
    out[1] = tot[1];
  }
}

// ---------------------------------------------------------------------------
// set-up kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T *x, T v, int n) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) x[q] = v;
}

__global__ void k_to_f32(const double *a, float *b, int n) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
    b[q] = (float)a[q];
}

// Galerkin coarse matrix for piecewise-constant prolongation, one thread per
// coarse cell; the fine block is visited in the order of
// additive_coarsen_system.cpp:103-219 (c summed, internal e/w per (j,k),
// n/s per (i,k), t/b per (i,j); external couplings summed over the faces).
// The fine matrix is symmetric 4-array; so is the coarse one.
__global__ void k_galerkin(Grid g, DMat<double> A, Grid gH, int CI, int CJ, int CK,
                           double *cH, double *wH, double *sH, double *bH) {
  for (int Q = blockIdx.x * blockDim.x + threadIdx.x; Q < gH.N; Q += gridDim.x * blockDim.x) {
    const int I = Q / gH.sI, remQ = Q - I * gH.sI, J = remQ / gH.nz, K = remQ - J * gH.nz;
    auto F = [&](int a, int b, int c) {
      return ((I * CI + a) * g.ny + (J * CJ + b)) * g.nz + (K * CK + c);
    };
    auto e_ = [&](int P) { return (P / g.sI) < g.nx - 1 ? A.w[P + g.sI] : 0.0; };
    auto n_ = [&](int P) { return ((P % g.sI) / g.nz) < g.ny - 1 ? A.s[P + g.nz] : 0.0; };
    auto t_ = [&](int P) { return (P % g.nz) < g.nz - 1 ? A.b[P + 1] : 0.0; };
    double c = 0.0;
    for (int a = 0; a < CI; a++)
      for (int b = 0; b < CJ; b++)
        for (int cc = 0; cc < CK; cc++) c += A.c[F(a, b, cc)];
    for (int b = 0; b < CJ; b++)
      for (int cc = 0; cc < CK; cc++) {
        for (int a = 0; a < CI - 1; a++) c -= e_(F(a, b, cc));
        for (int a = 1; a < CI; a++) c -= A.w[F(a, b, cc)];
      }
    double w = 0.0;
    for (int b = 0; b < CJ; b++)
      for (int cc = 0; cc < CK; cc++) w += A.w[F(0, b, cc)];
    for (int a = 0; a < CI; a++)
      for (int cc = 0; cc < CK; cc++) {
        for (int b = 0; b < CJ - 1; b++) c -= n_(F(a, b, cc));
        for (int b = 1; b < CJ; b++) c -= A.s[F(a, b, cc)];
      }
    double s = 0.0;
    for (int a = 0; a < CI; a++)
      for (int cc = 0; cc < CK; cc++) s += A.s[F(a, 0, cc)];
    for (int a = 0; a < CI; a++)
      for (int b = 0; b < CJ; b++) {
        for (int cc = 0; cc < CK - 1; cc++) c -= t_(F(a, b, cc));
        for (int cc = 1; cc < CK; cc++) c -= A.b[F(a, b, cc)];
      }
    double bb = 0.0;
    for (int a = 0; a < CI; a++)
      for (int b = 0; b < CJ; b++) bb += A.b[F(a, b, 0)];
    cH[Q] = c;
    wH[Q] = w;
    sH[Q] = s;
    bH[Q] = bb;
  }
}

// active flag of a coarse cell: OR over its block (additive_coarsen_flag.cpp)
__global__ void k_coarsen_flag(Grid g, const uint8_t *f, Grid gH, int CI, int CJ, int CK,
                               uint8_t *fH) {
  for (int Q = blockIdx.x * blockDim.x + threadIdx.x; Q < gH.N; Q += gridDim.x * blockDim.x) {
    const int I = Q / gH.sI, remQ = Q - I * gH.sI, J = remQ / gH.nz, K = remQ - J * gH.nz;
    int fl = 0;
    for (int a = 0; a < CI; a++)
      for (int b = 0; b < CJ; b++)
        for (int cc = 0; cc < CK; cc++)
          fl |= f[((I * CI + a) * g.ny + (J * CJ + b)) * g.nz + (K * CK + cc)];
    fH[Q] = (uint8_t)(fl != 0);
  }
}

}  // namespace psb
