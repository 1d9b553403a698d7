// psb_bkernels.h -- batched launches: one grid advances a step of many
// independent systems (own grids, matrices, hierarchies, solver states).
// Each block finds its system's segment (bmap), takes its block index inside
// the segment (seg_span) and runs the same kernel body as the single-system
// launch.  Per-system pointers come from device tables, so a captured graph
// stays valid when the caller's arrays change.
#pragma once

#include "psb_kernels.h"
#include "psb_march.h"

namespace psb {

// one level of one system's hierarchy (FP32 in the mixed-precision solver)
template <typename T>
struct BLev {
  Grid g;
  const T *c, *w, *s, *b;
  const T *d;         // flux-form diagonal (c minus the couplings)
  T *r;
  T *e[3];
  const uint8_t *af;
  int sI, sJ, sK;     // coarsening shift (0/1) per direction from the finer level
  int m;              // coarsest level: dense operator size (0 otherwise)
  const double *G;    // coarsest level: dense projected pseudo-inverse
  int TI, ntiles;     // TI > 0: the Jacobi sweeps (and on level 0 the CG p/q
                      // step) march, slabs of TI planes, ntiles tiles a plane
  int kcyc;           // this level's correction is the K-cycle's (in kz)
  T *kc, *kv, *kz;    // K-cycle: first cycle's output, A of it, the correction
  double *kst;        // K-cycle scalars rho1, alpha1, a, alpha2
  __device__ __forceinline__ DMat<T> mat() const {
    return DMat<T>{c, w, s, b, nullptr, nullptr, nullptr};
  }
  // the stencil kernels read the flux form: (d, w, s, b)
  __device__ __forceinline__ DMat<T> matflux() const {
    return DMat<T>{d, w, s, b, nullptr, nullptr, nullptr};
  }
};

// fine level of one system: the FP64 problem and the FP32 Krylov vectors
struct BFine {
  Grid g;
  DMat<double> A;        // symmetric 4-array FP64 matrix
  Metric met;
  const double *b;       // right-hand side (FP64)
  double *x;             // solution (FP64), updated at every reliable update
  float *d, *r, *p[2], *q;  // FP32 correction since the last update, CG vectors
  DevState *st;
};

#define PSB_BSEG_BEGIN                                    \
  const BSeg sg_ = segs[bmap[blockIdx.x]];                 \
  const BFine &F = fine[sg_.sys];                          \
  const Span sp = seg_span(sg_);

// damped Jacobi sweep on level `lev` (prolongation of levH's correction
// fused in when PROL: e[outH] times cfac, or where levH took the K-cycle its
// kz, already scaled; RHO on level 0's last post-sweep)
template <typename T, int V, bool PROL, bool RHO>
__global__ __launch_bounds__(kBlock) void kb_jacobi(const BSeg *segs, const uint16_t *bmap,
                                                    const BFine *fine, const BLev<T> *lev,
                                                    const BLev<T> *levH, int cur, int nxt,
                                                    int outH, T cfac, T omega, Work w,
                                                    bool flex) {
  PSB_BSEG_BEGIN
  if (F.st->done) return;
  const BLev<T> &L = lev[sg_.sys];
  // RHO with the flexible beta: the partials of z.q, q the last direction's A p
  const T *qf = RHO && flex ? (const T *)F.q : nullptr;
  ProlX<T, V> pr{};
  if (PROL) {
    const BLev<T> &H = levH[sg_.sys];
    pr.e = L.e[cur];
    pr.eH = H.kcyc ? H.kz : H.e[outH];
    pr.af = L.af;
    pr.cfac = H.kcyc ? T(1) : cfac;
    pr.sI = H.sI;
    pr.sJ = H.sJ;
    pr.sK = H.sK;
    pr.nYH = H.g.ny;
    pr.nZH = H.g.nz;
  }
  if constexpr (!PROL && V > 1) {
    if (L.TI > 0) {
      jacobi_march_body<T, V, false, RHO, T, true>(sp, L.g, L.matflux(), L.r, L.e[cur],
                                                   L.e[nxt], pr, omega, L.af, w, T(1), L.TI,
                                                   L.ntiles, qf);
      return;
    }
  }
  jacobi_body<T, V, true, PROL, RHO, T, true>(sp, L.g, L.matflux(), L.r, L.e[cur], L.e[nxt], pr,
                                              omega, L.af, w, T(1), qf);
}

// restriction of level lev into levH (all systems of the launch share the
// coarsening factors CI, CJ, CK), with levH's first sweep fused in
template <typename T, int V, int KV, int CI, int CJ, int CK>
__global__ __launch_bounds__(kBlock) void kb_restrict(const BSeg *segs, const uint16_t *bmap,
                                                      const BFine *fine, const BLev<T> *lev,
                                                      const BLev<T> *levH, int cur, T omega) {
  PSB_BSEG_BEGIN
  if (F.st->done) return;
  const BLev<T> &L = lev[sg_.sys];
  const BLev<T> &H = levH[sg_.sys];
  restrict_body<T, V, KV, true, CI, CJ, CK, T, true>(sp, L.g, L.matflux(), L.r, L.e[cur], L.af,
                                                     H.g, H.c, H.af, H.r, H.e[0], omega);
}

// K-cycle steps on level `lev` of the systems of the list (kc_av_body and
// the others, psb_kernels.h); each system's dot products summed over its own
// segment of blocks
template <typename T, int V>
__global__ __launch_bounds__(kBlock) void kb_kc_av(const BSeg *segs, const uint16_t *bmap,
                                                   const BFine *fine, const BLev<T> *lev,
                                                   int out, Work w) {
  PSB_BSEG_BEGIN
  if (F.st->done) return;
  const BLev<T> &L = lev[sg_.sys];
  kc_av_body<T, V, true>(sp, L.g, L.matflux(), T(1), L.e[out], L.r, L.kc, L.kv, w);
}

template <typename T, int V>
__global__ __launch_bounds__(kBlock) void kb_kc_r(const BSeg *segs, const uint16_t *bmap,
                                                  const BFine *fine, const BLev<T> *lev,
                                                  T omega, const double *partial) {
  PSB_BSEG_BEGIN
  if (F.st->done) return;
  const BLev<T> &L = lev[sg_.sys];
  kc_r_body<T, V>(sp, L.g, L.c, L.kv, L.r, L.e[0], omega, L.kst, sg_.nb,
                  partial + (size_t)sg_.b0 * 2);
}

template <typename T, int V>
__global__ __launch_bounds__(kBlock) void kb_kc_dw(const BSeg *segs, const uint16_t *bmap,
                                                   const BFine *fine, const BLev<T> *lev,
                                                   int out, Work w) {
  PSB_BSEG_BEGIN
  if (F.st->done) return;
  const BLev<T> &L = lev[sg_.sys];
  kc_dw_body<T, V, true>(sp, L.g, L.matflux(), T(1), L.e[out], L.kv, L.r, w);
}

template <typename T, int V>
__global__ __launch_bounds__(kBlock) void kb_kc_comb(const BSeg *segs, const uint16_t *bmap,
                                                     const BFine *fine, const BLev<T> *lev,
                                                     int out, const double *partial) {
  PSB_BSEG_BEGIN
  if (F.st->done) return;
  const BLev<T> &L = lev[sg_.sys];
  kc_comb_body<T, V>(sp, L.g, L.kc, L.e[out], L.kz, L.kst, sg_.nb,
                     partial + (size_t)sg_.b0 * 3);
}

// coarsest level: e[out] = G r
template <typename T>
__global__ __launch_bounds__(kBlock) void kb_gemv(const BSeg *segs, const uint16_t *bmap,
                                                  const BFine *fine, const BLev<T> *lev,
                                                  int out) {
  PSB_BSEG_BEGIN
  if (F.st->done) return;
  const BLev<T> &L = lev[sg_.sys];
  gemv_body<T>(sp.bl, L.m, L.G, L.r, L.e[out]);
}

// CG p/q step on the fine level, z = level 0's e[zi]
template <int V>
__global__ __launch_bounds__(kBlock) void kb_cg_pq(const BSeg *segs, const uint16_t *bmap,
                                                   const BFine *fine, const BLev<float> *lev0,
                                                   int zi, int par, Work w) {
  PSB_BSEG_BEGIN
  if (F.st->done) return;
  const BLev<float> &L = lev0[sg_.sys];
  if constexpr (V > 1) {
    if (L.TI > 0) {
      cg_pq_march_body<float, V, true, true>(sp, F.g, L.matflux(), L.e[zi], F.p[par],
                                             F.p[par ^ 1], F.q, L.af, F.st, w, L.TI, L.ntiles);
      return;
    }
  }
  cg_pq_body<float, V, true, false, true>(sp, F.g, L.matflux(), L.e[zi], F.p[par], F.p[par ^ 1],
                                          F.q, L.af, F.st, w);
}

// CG x/r step (x: the FP32 correction d), first sweep of the next V-cycle
template <int V>
__global__ __launch_bounds__(kBlock) void kb_cg_xr(const BSeg *segs, const uint16_t *bmap,
                                                   const BFine *fine, const BLev<float> *lev0,
                                                   int par, float omega, Work w) {
  PSB_BSEG_BEGIN
  if (F.st->done) return;
  const BLev<float> &L = lev0[sg_.sys];
  cg_xr_body<float, V, false>(sp, F.g, L.c, F.met, F.p[par ^ 1], F.q, F.d, F.r, L.e[0], omega,
                              F.st, w);
}

// reliable update: x += d (gated on a pending update unless starting)
template <int V>
__global__ __launch_bounds__(kBlock) void kb_ru_xupdate(const BSeg *segs, const uint16_t *bmap,
                                                        const BFine *fine, bool gated) {
  PSB_BSEG_BEGIN
  if (gated && !F.st->update) return;
  ru_xupdate_body<V>(sp, F.g, F.x, F.d);
}

// reliable update: r32 = (float)(b - A x) in FP64, e1 = omega r32 / c32
template <int V>
__global__ __launch_bounds__(kBlock) void kb_ru_resid(const BSeg *segs, const uint16_t *bmap,
                                                      const BFine *fine, const BLev<float> *lev0,
                                                      float omega, bool gated, Work w) {
  PSB_BSEG_BEGIN
  if (gated && !F.st->update) return;
  const BLev<float> &L = lev0[sg_.sys];
  ru_resid_body<V, true>(sp, F.g, F.A, F.met, F.b, F.x, L.c, F.r, L.e[0], omega, w,
                         F.st->norm == 0);
}

// zero the FP32 correction and both search directions
template <int V>
__global__ __launch_bounds__(kBlock) void kb_zero(const BSeg *segs, const uint16_t *bmap,
                                                  const BFine *fine) {
  PSB_BSEG_BEGIN
  const VecN<float, V> z = zerov<float, V>();
  PSB_SPAN_BEGIN(F.g, V, sp)
  stv<float, V>(F.d + P, z);
  stv<float, V>(F.p[0] + P, z);
  stv<float, V>(F.p[1] + P, z);
  (void)i; (void)j; (void)k;
  PSB_TILE_END
}

// pass 2 of the reductions for every system: block s sums the partials of
// system s's fine-level segment fsegs[s] and applies op to its state;
// flags[s] mirrors the stop flag for the host
template <int NR>
__global__ __launch_bounds__(kBlock) void kb_finalize(int op, const BSeg *fsegs,
                                                      const BFine *fine, const double *partial,
                                                      int *flags) {
  const int sys = blockIdx.x;
  const BFine &F = fine[sys];
  DevState *st = F.st;
  if (!fin_skip(op, st)) {
    double acc[NR];
    fin_sum<NR>(fsegs[sys].b0, fsegs[sys].nb, partial, acc);
    if (threadIdx.x == 0) fin_apply(op, acc, st, nullptr, nullptr);
  }
  if (threadIdx.x == 0) flags[sys] = st->done;
}

}  // namespace psb
