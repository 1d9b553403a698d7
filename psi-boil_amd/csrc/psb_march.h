// psb_march.h -- stencil kernels that march through the i planes.
//
// A block owns a tile of kBlock cell vectors of one (j, k) plane and walks
// TI consecutive planes; each thread keeps its own column's operand at
// planes i-1, i, i+1 in registers, so the i-neighbours are loaded once
// instead of three times and every plane of the operand streams from HBM
// once per tile (the j and k neighbours come from the L1 / L2 copies the
// neighbouring threads and tiles load).  The arithmetic is that of the
// span kernels (same gather, same order): results are bit-identical.
#pragma once

#include "psb_kernels.h"

namespace psb {

// the operand with the thread's own column at planes i-1 / i / i+1 held in
// registers; other points read through the base operand
template <typename T, int V, class B>
struct MarchX {
  B base;
  VecN<T, V> prev, cur, next;
  int i, j;
  __device__ __forceinline__ VecN<T, V> vec(int P, int ii, int jj, int kk) const {
    if (ii < i) return prev;
    if (ii > i) return next;
    if (jj == j) return cur;
    return base.vec(P, ii, jj, kk);
  }
  __device__ __forceinline__ T at(int Q, int ii, int jj, int kk) const {
    return base.at(Q, ii, jj, kk);
  }
};
template <typename T, int V, class B>
__device__ __forceinline__ VecN<T, V> x_nbj(const MarchX<T, V, B> &x, int P, int ii, int jj,
                                            int kk) {
  return x.base.vec(P, ii, jj, kk);
}
template <typename T, int V, class B>
__device__ __forceinline__ VecN<T, V> x_nbi(const MarchX<T, V, B> &x, bool up, int, int, int,
                                            int) {
  return up ? x.next : x.prev;
}

// launch geometry: tiles of kBlock vectors per plane, slabs of TI planes,
// about kMaxBlocks blocks in all
struct MarchShape {
  int TI, ntiles, nblocks;
};
#ifndef PSB_MARCH_BLOCKS
#define PSB_MARCH_BLOCKS 1024
#endif
// target grid (<= kMaxBlocks): 1024 blocks, i.e. slabs of 16 planes at
// 256^3, measured best against 512 and 2048 (27.1 / 29.4 / 27.8 ms per mixed
// solve, 55.0 / 56.3 / 56.3 ms per FP64 solve)
constexpr int kMarchBlocks = PSB_MARCH_BLOCKS;
// nplanes planes of nvp thread items each
inline MarchShape march_shape_n(int nplanes, long long nvp) {
  MarchShape m;
  m.ntiles = (int)((nvp + kBlock - 1) / kBlock);
  long long TI = ((long long)nplanes * m.ntiles + kMarchBlocks - 1) / kMarchBlocks;
  if (TI < 1) TI = 1;
  // the reductions keep one partial per block: at most kMaxBlocks blocks
  while (TI < nplanes && m.ntiles * ((nplanes + TI - 1) / TI) > kMaxBlocks) TI++;
  m.TI = (int)TI;
  m.nblocks = m.ntiles * (int)((nplanes + TI - 1) / TI);
  return m;
}
template <int V>
inline MarchShape march_shape(const Grid &g) {
  return march_shape_n(g.nx, (long long)g.ny * (g.nz / V));
}
// the march applies when the vectors tile the rows and the grid fits the
// reduction buffer (a plane of more than kMaxBlocks tiles does not)
template <int V>
inline bool march_ok(const Grid &g, MarchShape &m) {
  if (g.nz % V != 0) return false;
  m = march_shape<V>(g);
  return m.nblocks <= kMaxBlocks;
}

// the (j, k) of this thread's vector and its slab; act_ false past the
// plane's last vector (such threads skip the march but stay for the
// block's reductions)
#define PSB_MARCH_BEGIN(g, V, TI, ntiles)                                     \
  const int wi_ = xcd_block();                                                \
  const int slab_ = wi_ / (ntiles), tile_ = wi_ - slab_ * (ntiles);           \
  const int nv_ = (g).nz / (V);                                               \
  const int q_ = tile_ * kBlock + (int)threadIdx.x;                          \
  const bool act_ = q_ < (g).ny * nv_;                                        \
  const int j = act_ ? q_ / nv_ : 0, k = act_ ? (q_ - j * nv_) * (V) : 0;     \
  const int i0_ = slab_ * (TI), i1_ = act_ ? min(i0_ + (TI), (g).nx) : i0_;

// y = A x
template <typename T, int V, bool SYM>
__global__ __launch_bounds__(kBlock) void k_spmv_march(Grid g, DMat<T> A, const T *x, T *y,
                                                       int TI, int ntiles) {
  PSB_MARCH_BEGIN(g, V, TI, ntiles)
  if (!act_) return;
  // software-pipelined as the Jacobi sweep: plane i + 1's coefficients and
  // plane i + 2's operand are loaded while plane i is computed
  int P = (i0_ * g.ny + j) * g.nz + k;
  VecN<T, V> prev = i0_ > 0 ? ldv<T, V>(x + P - g.sI) : zerov<T, V>();
  VecN<T, V> cur = ldv<T, V>(x + P);
  const bool up0 = i0_ + 1 < g.nx;
  VecN<T, V> next = vsel<T, V>(up0, ldv<T, V>(x + (up0 ? P + g.sI : P)));
  Coef<T, V, SYM> cf = load_coef<T, V, SYM>(g, A, P, i0_, j, k);
  for (int i = i0_; i < i1_; i++, P += g.sI) {
    const bool in1 = i + 1 < i1_, up2 = in1 && i + 2 < g.nx;
    const VecN<T, V> nn = vsel<T, V>(up2, ldv<T, V>(x + (up2 ? P + 2 * g.sI : P)));
    const Coef<T, V, SYM> cfn =
        load_coef<T, V, SYM>(g, A, in1 ? P + g.sI : P, in1 ? i + 1 : i, j, k);
    MarchX<T, V, ArrX<T, V>> X{ArrX<T, V>{x}, prev, cur, next, i, j};
    Gath<T, V> G = gather_c<T, V, SYM>(g, cf, X, P, i, j, k);
    stv<T, V>(y + P, apply_A<T, V>(G));
    prev = cur;
    cur = next;
    next = nn;
    cf = cfn;
  }
}

// coefficients loaded a plane ahead, except in the last post-sweep (RHO:
// measured the same either way at 256^3, and without it the kernel holds
// fewer registers)
template <bool RHO>
constexpr bool kPipe = !RHO;

// damped Jacobi sweep e_out = e + omega (r - A e) / c, the operand e or,
// with PROL, e + cfac P eH (the first post-sweep); RHO: the last post-sweep
// of level 0 (output times zscale, partials r.z, sum z, sum r, active
// cells per block).  The same arithmetic as k_jacobi.
// one item: tile `tile` of slab `slab` (TI planes)
template <typename T, int V, bool PROL, bool RHO, typename M, bool FLUX>
__device__ __forceinline__ void jacobi_march_item(const Grid &g, const DMat<M> &A, const T *r,
                                                  const T *e, T *eout, const ProlX<T, V> &pr,
                                                  T omega, const uint8_t *af, T zscale, int TI,
                                                  int slab, int tile, const T *qf,
                                                  double (&acc)[kRhoParts]) {
  const int nv = g.nz / V;
  const int q = tile * kBlock + (int)threadIdx.x;
  if (q >= g.ny * nv) return;
  const int j = q / nv, k = (q - j * nv) * V;
  const int i0 = slab * TI, i1 = min(i0 + TI, g.nx);
  auto opv = [&](int Q, int ii) {
    if constexpr (PROL) return pr.vec(Q, ii, j, k);
    else return ldv<T, V>(e + Q);
  };
  // software-pipelined: plane i + 1's coefficients and plane i + 2's
  // operand are loaded while plane i is computed (past the slab's end the
  // loads read plane i again and are discarded)
  int P = (i0 * g.ny + j) * g.nz + k;
  VecN<T, V> prev = i0 > 0 ? opv(P - g.sI, i0 - 1) : zerov<T, V>();
  VecN<T, V> cur = opv(P, i0);
  const bool up0 = i0 + 1 < g.nx;
  VecN<T, V> next = vsel<T, V>(up0, opv(up0 ? P + g.sI : P, up0 ? i0 + 1 : i0));
  Coef<M, V, true> cf;
  if constexpr (kPipe<RHO>) cf = load_coef<M, V, true>(g, A, P, i0, j, k);
  for (int i = i0; i < i1; i++, P += g.sI) {
    const bool in1 = i + 1 < i1, up2 = in1 && i + 2 < g.nx;
    const VecN<T, V> nn = vsel<T, V>(up2, opv(up2 ? P + 2 * g.sI : P, up2 ? i + 2 : i));
    Coef<M, V, true> cfn;
    if constexpr (kPipe<RHO>) cfn = load_coef<M, V, true>(g, A, in1 ? P + g.sI : P, in1 ? i + 1 : i, j, k);
    else cf = load_coef<M, V, true>(g, A, P, i, j, k);
    // (without q, r is read twice rather than under a branch)
    const VecN<T, V> R = ldv<T, V>(r + P);
    VecN<T, V> Qf;
    if constexpr (RHO) Qf = ldv<T, V>((qf ? qf : r) + P);
    Gath<T, V> G;
    if constexpr (PROL) {
      MarchX<T, V, ProlX<T, V>> X{pr, prev, cur, next, i, j};
      G = gather_c<T, V, true, FLUX>(g, cf, X, P, i, j, k);
    } else {
      MarchX<T, V, ArrX<T, V>> X{ArrX<T, V>{e}, prev, cur, next, i, j};
      G = gather_c<T, V, true, FLUX>(g, cf, X, P, i, j, k);
    }
    const VecN<T, V> t = resid<T, V>(R, G);
    VecN<T, V> E;
#pragma unroll
    for (int u = 0; u < V; u++) {
      E.v[u] = G.x.v[u] + omega * t.v[u] / G.c.v[u];
      if (RHO) {
        E.v[u] = zscale * E.v[u];
        rho_accq<T>(acc, R.v[u], E.v[u], qf != nullptr, Qf.v[u], !af || af[P + u]);
      }
    }
    stv<T, V>(eout + P, E);
    prev = cur;
    cur = next;
    next = nn;
    if constexpr (kPipe<RHO>) cf = cfn;
  }
}

// all items of a grid of ntiles x slabs, block bl of nb taking items
// bl, bl + nb, ... (one item per block in the single-system launch)
template <typename T, int V, bool PROL, bool RHO, typename M, bool FLUX>
__device__ __forceinline__ void jacobi_march_body(Span sp, const Grid &g, const DMat<M> &A,
                                                  const T *r, const T *e, T *eout,
                                                  const ProlX<T, V> &pr, T omega,
                                                  const uint8_t *af, Work w, T zscale, int TI,
                                                  int ntiles, const T *qf = nullptr) {
  double acc[kRhoParts] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int items = ntiles * ((g.nx + TI - 1) / TI);
  for (int wi = sp.bl; wi < items; wi += sp.nb)
    jacobi_march_item<T, V, PROL, RHO, M, FLUX>(g, A, r, e, eout, pr, omega, af, zscale, TI,
                                                wi / ntiles, wi % ntiles, qf, acc);
  if (RHO) block_partial<kRhoParts>(acc, w.partial);
}

// damped Jacobi sweep e_out = e + omega (r - A e) / c, the operand e or,
// with PROL, e + cfac P eH (the first post-sweep); RHO: the last post-sweep
// of level 0 (output times zscale, partials r.z, sum z, sum r, active
// cells per block).  The same arithmetic as k_jacobi.
template <typename T, int V, bool PROL, bool RHO, typename M, bool FLUX>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_jacobi_march(Grid g, DMat<M> A, const T *r,
                                                         const T *e, T *eout, ProlX<T, V> pr,
                                                         T omega, const uint8_t *af,
                                                         const DevState *st, Work w, T zscale,
                                                         int TI, int ntiles, const T *qf) {
  if (skip_done(st, g.N)) return;
  jacobi_march_body<T, V, PROL, RHO, M, FLUX>(grid_span(), g, A, r, e, eout, pr, omega, af, w,
                                              zscale, TI, ntiles, qf);
}

// CG step 1 (cg.cpp:75-108): p = z + beta p_old with z = r/c (DIAG) or the
// projected preconditioner output, q = A p, partial p.q -- as k_cg_pq
// CG step 1 item (cg.cpp:75-108): p = z + beta p_old, q = A p, p.q
template <typename T, int V, bool SYM, bool FLUX, typename M, class B>
__device__ __forceinline__ void cg_pq_march_item(const Grid &g, const DMat<M> &A, const B &base,
                                                 T *pnew, T *q, T qscale, int TI, int slab,
                                                 int tile, double (&acc)[1]) {
  const int nv = g.nz / V;
  const int qi = tile * kBlock + (int)threadIdx.x;
  if (qi >= g.ny * nv) return;
  const int j = qi / nv, k = (qi - j * nv) * V;
  const int i0 = slab * TI, i1 = min(i0 + TI, g.nx);
  int P = (i0 * g.ny + j) * g.nz + k;
  VecN<T, V> prev = i0 > 0 ? base.vec(P - g.sI, i0 - 1, j, k) : zerov<T, V>();
  VecN<T, V> cur = base.vec(P, i0, j, k);
  for (int i = i0; i < i1; i++, P += g.sI) {
    const bool up = i + 1 < g.nx;
    const VecN<T, V> next = vsel<T, V>(up, base.vec(up ? P + g.sI : P, up ? i + 1 : i, j, k));
    MarchX<T, V, B> X{base, prev, cur, next, i, j};
    Gath<T, V> G = gather<T, V, SYM, FLUX>(g, A, X, P, i, j, k);
    VecN<T, V> Q = apply_A<T, V>(G);
    if constexpr (!std::is_same<T, M>::value) {
#pragma unroll
      for (int u = 0; u < V; u++) Q.v[u] = Q.v[u] * qscale;
    }
    stv<T, V>(pnew + P, G.x);
    stv<T, V>(q + P, Q);
#pragma unroll
    for (int u = 0; u < V; u++) acc[0] += (double)(G.x.v[u] * Q.v[u]);
    prev = cur;
    cur = next;
  }
}

// the projected-z form (the preconditioned CG), over the items of block bl
template <typename T, int V, bool SYM, bool FLUX, typename M = T>
__device__ __forceinline__ void cg_pq_march_body(Span sp, const Grid &g, const DMat<M> &A,
                                                 const T *z, const T *pold, T *pnew, T *q,
                                                 const uint8_t *af, const DevState *st, Work w,
                                                 int TI, int ntiles, T qscale = T(1)) {
  double acc[1] = {0.0};
  const ZPX<T, V> base{z, pold, af, (T)st->beta, (T)st->zmean};
  const int items = ntiles * ((g.nx + TI - 1) / TI);
  for (int wi = sp.bl; wi < items; wi += sp.nb)
    cg_pq_march_item<T, V, SYM, FLUX, M>(g, A, base, pnew, q, qscale, TI, wi / ntiles,
                                         wi % ntiles, acc);
  block_partial<1>(acc, w.partial);
}

// CG step 1 (as k_cg_pq): z = r/c (DIAG) or the projected preconditioner
// output; M = _Float16: the level's FP16 copy scale * (d, w, s, b) (flux
// form, not with DIAG), q = the product times qscale = 1 / scale
template <typename T, int V, bool SYM, bool DIAG, bool FLUX, typename M = T>
__global__ __launch_bounds__(kBlock) void k_cg_pq_march(Grid g, DMat<M> A, const T *r_or_z,
                                                        const T *pold, T *pnew, T *q,
                                                        const uint8_t *af, DevState *st,
                                                        Work w, int TI, int ntiles,
                                                        T qscale = T(1)) {
  static_assert(!DIAG || std::is_same<T, M>::value, "the diagonal operand divides by A.c");
  if (st->done) return;
  if constexpr (DIAG) {
    double acc[1] = {0.0};
    const DiagPX<T, V> base{r_or_z, (const T *)A.c, pold, (T)st->beta};
    const Span sp = grid_span();
    const int items = ntiles * ((g.nx + TI - 1) / TI);
    for (int wi = sp.bl; wi < items; wi += sp.nb)
      cg_pq_march_item<T, V, SYM, FLUX, M>(g, A, base, pnew, q, qscale, TI, wi / ntiles,
                                           wi % ntiles, acc);
    block_partial<1>(acc, w.partial);
  } else {
    cg_pq_march_body<T, V, SYM, FLUX, M>(grid_span(), g, A, r_or_z, pold, pnew, q, af, st, w,
                                         TI, ntiles, qscale);
  }
}

// restriction (as k_restrict, CI = 2, one vector of V fine cells along k per
// thread and row): a thread owns a coarse (J, k-vector) column and walks a
// slab of coarse planes; per coarse plane it needs the operand at fine
// planes 2I-1 .. 2I+2 of its CJ rows, of which 2I-1 and 2I are the previous
// step's 2I+1 and 2I+2.  The sums run in the same order as k_restrict.
template <typename T, int V, bool SYM, int CJ, int CK, typename M, bool FLUX>
__global__ __launch_bounds__(kBlock) void k_restrict_march(Grid g, DMat<M> A, const T *r,
                                                           const T *e, const uint8_t *af,
                                                           Grid gH, const T *cH,
                                                           const uint8_t *afH, T *rH, T *eH1,
                                                           T omega, DevState *st, int TI,
                                                           int ntiles) {
  if (skip_done(st, g.N)) return;
  constexpr int NC = V / CK;  // coarse cells per thread
  const int wi = xcd_block();
  const int slab = wi / ntiles, tile = wi - slab * ntiles;
  const int nkv = g.nz / V;
  const int q = tile * kBlock + (int)threadIdx.x;
  if (q >= gH.ny * nkv) return;
  const int J = q / nkv, kv = q - J * nkv;
  const int k = kv * V;
  const int I0 = slab * TI, I1 = min(I0 + TI, gH.nx);
  ArrX<T, V> X{e};
  VecN<T, V> ca[CJ], cb[CJ];  // operand at fine planes 2I-1 and 2I, per row
#pragma unroll
  for (int bb = 0; bb < CJ; bb++) {
    const int j = J * CJ + bb;
    const int P = (2 * I0 * g.ny + j) * g.nz + k;
    ca[bb] = I0 > 0 ? ldv<T, V>(e + P - g.sI) : zerov<T, V>();
    cb[bb] = ldv<T, V>(e + P);
  }
  for (int I = I0; I < I1; I++) {
    VecN<T, V> cc[CJ], cd[CJ];  // fine planes 2I+1 and 2I+2 (past the last: zero)
    const bool up = 2 * I + 2 < g.nx;
#pragma unroll
    for (int bb = 0; bb < CJ; bb++) {
      const int j = J * CJ + bb;
      const int P = ((2 * I + 1) * g.ny + j) * g.nz + k;
      cc[bb] = ldv<T, V>(e + P);
      cd[bb] = vsel<T, V>(up, ldv<T, V>(e + (up ? P + g.sI : P)));
    }
    T s[NC];
#pragma unroll
    for (int u = 0; u < NC; u++) s[u] = T(0);
#pragma unroll
    for (int a = 0; a < 2; a++) {
#pragma unroll
      for (int bb = 0; bb < CJ; bb++) {
        const int i = 2 * I + a, j = J * CJ + bb;
        const int P = (i * g.ny + j) * g.nz + k;
        MarchX<T, V, ArrX<T, V>> Xm{X, a == 0 ? ca[bb] : cb[bb], a == 0 ? cb[bb] : cc[bb],
                                    a == 0 ? cc[bb] : cd[bb], i, j};
        Gath<T, V> G = gather<T, V, SYM, FLUX>(g, A, Xm, P, i, j, k);
        VecN<T, V> R = resid<T, V>(ldv<T, V>(r + P), G);
        T t[V];
#pragma unroll
        for (int u = 0; u < V; u++) t[u] = (!af || af[P + u]) ? R.v[u] : T(0);
#pragma unroll
        for (int u = 0; u < NC; u++)
#pragma unroll
          for (int c2 = 0; c2 < CK; c2++) s[u] += t[u * CK + c2];
      }
    }
    const int Q0 = (I * gH.ny + J) * gH.nz + kv * NC;
#pragma unroll
    for (int u = 0; u < NC; u++) {
      const bool act = !afH || afH[Q0 + u];
      const T v = act ? s[u] : T(0);
      rH[Q0 + u] = v;
      eH1[Q0 + u] = omega * v / cH[Q0 + u];
    }
#pragma unroll
    for (int bb = 0; bb < CJ; bb++) {
      ca[bb] = cc[bb];
      cb[bb] = cd[bb];
    }
  }
}

}  // namespace psb
