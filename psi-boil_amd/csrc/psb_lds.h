// psb_lds.h -- marching stencil kernels with the operand's planes staged in
// LDS.
//
// The tiling is that of psb_march.h (a block owns R whole rows of a (j, k)
// plane and walks TI consecutive i planes; the operand's own column at
// planes i-1 / i / i+1 in registers), but the operand of the current plane
// -- the block's R rows plus the row below and the row above -- is written
// to LDS once per plane, and the k-1 / k+V and j-1 / j+1 neighbours are read
// from there instead of through the vector memory path.  The operand is
// formed on the fly from its arrays (e, or e + cfac P eH for the first
// post-sweep, or z + beta p for the CG step) once per cell and plane
// (the halo rows once more, by the block that needs them), where the
// register-marching kernels form it again for every neighbour that reads it.
// The next plane's loads are issued before the current plane is computed.
// Same gather, same arithmetic: results are bit-identical to k_jacobi,
// k_jacobi_march, k_cg_pq and k_cg_pq_march.
#pragma once

#include "psb_march.h"

namespace psb {

// whether the LDS tile applies: a row is nz / V thread vectors and a block
// holds R >= 2 whole rows (so that 2 * nz / V threads load the halo rows)
template <int V>
__host__ __device__ inline bool lds_ok(const Grid &g) {
  if (g.nz % V != 0) return false;
  const int nv = g.nz / V;
  return kBlock % nv == 0 && kBlock / nv >= 2;
}

// the operand at the staged plane: own column at i-1 / i / i+1 from
// registers, every other point of plane i from the LDS rows
template <typename T, int V>
struct LdsX {
  const T *rows;  // plane i: row jj at (jj - j0 + 1) * nz
  VecN<T, V> prev, cur, next;
  int i, j, j0, nz;
  __device__ __forceinline__ VecN<T, V> vec(int, int ii, int jj, int kk) const {
    if (ii < i) return prev;
    if (ii > i) return next;
    if (jj == j) return cur;
    return *reinterpret_cast<const VecN<T, V> *>(rows + (jj - j0 + 1) * nz + kk);
  }
  __device__ __forceinline__ T at(int, int, int jj, int kk) const {
    return rows[(jj - j0 + 1) * nz + kk];
  }
};
template <typename T, int V>
__device__ __forceinline__ VecN<T, V> x_nbj(const LdsX<T, V> &x, int, int, int jj, int kk) {
  return *reinterpret_cast<const VecN<T, V> *>(x.rows + (jj - x.j0 + 1) * x.nz + kk);
}
template <typename T, int V>
__device__ __forceinline__ VecN<T, V> x_nbi(const LdsX<T, V> &x, bool up, int, int, int, int) {
  return up ? x.next : x.prev;
}

// one (slab, tile) item: stage, gather and epi(P, i, j, k, G, pre(P)) per
// plane.  Every thread of the block runs the same number of planes (the LDS
// barriers); threads past the grid's last row stage nothing.  A plane's
// global loads are issued without branches (lanes with nothing to load
// read row 0 and discard it): the next plane's operand before the barrier,
// the coefficients and what pre(P) reads for the epilogue before it too
// (PRE) or after it.  Measured per kernel at 256^3: the first post-sweep
// (its operand formed from two arrays) gains from PRE, the CG step and the
// FP64 sweep do not.
template <bool PRE, typename T, int V, bool SYM, bool FLUX, typename M, class B, class Pre,
          class Epi>
__device__ __forceinline__ void lds_march_item(const Grid &g, const DMat<M> &A, const B &base,
                                               int TI, int slab, int tile, T (*buf)[2 * kBlock * V],
                                               Pre pre, Epi epi) {
  const int nv = g.nz / V, R = kBlock / nv;
  const int t = (int)threadIdx.x;
  const int j0 = tile * R;
  const int r = t / nv, kv = t - r * nv;
  const int j = j0 + r, k = kv * V;
  const bool act = j < g.ny;
  // halo: threads [0, nv) the row below the tile, [nv, 2 nv) the row above
  const int hr = t < nv ? -1 : R;
  const int hj = j0 + hr, hk = (t < nv ? t : t - nv) * V;
  const bool hact = t < 2 * nv && hj >= 0 && hj < g.ny;
  const int ja = act ? j : 0, hja = hact ? hj : 0, hka = hact ? hk : 0;
  const int i0 = slab * TI, i1 = min(i0 + TI, g.nx);
  int P = (i0 * g.ny + ja) * g.nz + k;
  int HP = (i0 * g.ny + hja) * g.nz + hka;
  VecN<T, V> prev = zerov<T, V>();
  if (i0 > 0) prev = base.vec(P - g.sI, i0 - 1, ja, k);
  VecN<T, V> cur = base.vec(P, i0, ja, k), hcur = base.vec(HP, i0, hja, hka);
  for (int i = i0; i < i1; i++, P += g.sI, HP += g.sI) {
    const bool up = i + 1 < g.nx;
    const int iu = up ? i + 1 : i;
    const VecN<T, V> next = vsel<T, V>(up, base.vec(up ? P + g.sI : P, iu, ja, k));
    const VecN<T, V> hnext = base.vec(up ? HP + g.sI : HP, iu, hja, hka);
    Coef<M, V, SYM> cf;
    decltype(pre(P)) pv;
    if constexpr (PRE) {
      cf = load_coef<M, V, SYM>(g, A, P, i, ja, k);
      pv = pre(P);
    }
    T *rows = buf[i & 1];
    if (act) *reinterpret_cast<VecN<T, V> *>(rows + (r + 1) * g.nz + k) = cur;
    if (hact) *reinterpret_cast<VecN<T, V> *>(rows + (hr + 1) * g.nz + hk) = hcur;
    __syncthreads();
    if constexpr (!PRE) {
      cf = load_coef<M, V, SYM>(g, A, P, i, ja, k);
      pv = pre(P);
    }
    if (act) {
      const LdsX<T, V> X{rows, prev, cur, next, i, j, j0, g.nz};
      const Gath<T, V> G = gather_c<T, V, SYM, FLUX>(g, cf, X, P, i, j, k);
      epi(P, i, j, k, G, pv);
    }
    prev = cur;
    cur = next;
    hcur = hnext;
  }
}

struct NoPre {};

// damped Jacobi sweep e_out = e + omega (r - A e) / c (the operand e, or
// with PROL e + cfac P eH); RHO: the last post-sweep of level 0 (output
// times zscale, the rho partials).  As k_jacobi_march.
template <typename T, int V, bool PROL, bool RHO, typename M, bool FLUX>
__global__ __launch_bounds__(kBlock) void k_jacobi_lds(Grid g, DMat<M> A, const T *r,
                                                       const T *e, T *eout, ProlX<T, V> pr,
                                                       T omega, const uint8_t *af,
                                                       const DevState *st, Work w, T zscale,
                                                       int TI, int ntiles, const T *qf) {
  if (skip_done(st, g.N)) return;
  __shared__ T buf[2][2 * kBlock * V];
  double acc[kRhoParts] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int wi = xcd_block();
  const int slab = wi / ntiles, tile = wi - slab * ntiles;
  struct RQ {
    VecN<T, V> R, Q;
  };
  // r, and q for the rho partials (without q, r read twice)
  auto pre = [&](int P) {
    RQ o;
    o.R = ldv<T, V>(r + P);
    if constexpr (RHO) o.Q = ldv<T, V>((qf ? qf : r) + P);
    return o;
  };
  auto epi = [&](int P, int, int, int, const Gath<T, V> &G, const RQ &pv) {
    const VecN<T, V> t = resid<T, V>(pv.R, G);
    VecN<T, V> E;
#pragma unroll
    for (int u = 0; u < V; u++) {
      E.v[u] = G.x.v[u] + omega * t.v[u] / G.c.v[u];
      if (RHO) {
        E.v[u] = zscale * E.v[u];
        rho_accq<T>(acc, pv.R.v[u], E.v[u], qf != nullptr, pv.Q.v[u], !af || af[P + u]);
      }
    }
    stv<T, V>(eout + P, E);
  };
  if constexpr (PROL)
    lds_march_item<true, T, V, true, FLUX, M>(g, A, pr, TI, slab, tile, buf, pre, epi);
  else
    lds_march_item<false, T, V, true, FLUX, M>(g, A, ArrX<T, V>{e}, TI, slab, tile, buf, pre,
                                               epi);
  if (RHO) block_partial<kRhoParts>(acc, w.partial);
}

// CG step 1 (cg.cpp:75-108): p = z + beta p_old (DIAG: z = r / c; else the
// projected preconditioner output), q = A p, partial p.q.  As k_cg_pq_march.
template <typename T, int V, bool SYM, bool DIAG, bool FLUX, typename M = T>
__global__ __launch_bounds__(kBlock) void k_cg_pq_lds(Grid g, DMat<M> A, const T *r_or_z,
                                                      const T *pold, T *pnew, T *q,
                                                      const uint8_t *af, DevState *st, Work w,
                                                      int TI, int ntiles, T qscale) {
  static_assert(!DIAG || std::is_same<T, M>::value, "the diagonal operand divides by A.c");
  if (st->done) return;
  __shared__ T buf[2][2 * kBlock * V];
  double acc[1] = {0.0};
  const int wi = xcd_block();
  const int slab = wi / ntiles, tile = wi - slab * ntiles;
  auto pre = [](int) { return NoPre{}; };
  auto epi = [&](int P, int, int, int, const Gath<T, V> &G, NoPre) {
    VecN<T, V> Q = apply_A<T, V>(G);
    if constexpr (!std::is_same<T, M>::value) {
#pragma unroll
      for (int u = 0; u < V; u++) Q.v[u] = Q.v[u] * qscale;
    }
    stv<T, V>(pnew + P, G.x);
    stv<T, V>(q + P, Q);
#pragma unroll
    for (int u = 0; u < V; u++) acc[0] += (double)(G.x.v[u] * Q.v[u]);
  };
  if constexpr (DIAG) {
    const DiagPX<T, V> base{r_or_z, (const T *)A.c, pold, (T)st->beta};
    lds_march_item<false, T, V, SYM, FLUX, M>(g, A, base, TI, slab, tile, buf, pre, epi);
  } else {
    const ZPX<T, V> base{r_or_z, pold, af, (T)st->beta, (T)st->zmean};
    lds_march_item<false, T, V, SYM, FLUX, M>(g, A, base, TI, slab, tile, buf, pre, epi);
  }
  block_partial<1>(acc, w.partial);
}

// reliable update of the mixed solve: r32 = (float)(b - A x) in FP64, e1 =
// omega r32 / c32; partials r64^2, r64^2 / dV^2 (vd), r32^2, b^2.  As
// k_ru_resid, marching with x staged in LDS.
template <int V, bool SYM>
__global__ __launch_bounds__(kBlock) void k_ru_resid_lds(Grid g, DMat<double> A, Metric m,
                                                         const double *b, const double *x,
                                                         const float *c32, float *r32,
                                                         float *e1, float omega,
                                                         const DevState *st, bool gated,
                                                         Work w, int TI, int ntiles) {
  if (gated && !st->update) return;
  __shared__ double buf[2][2 * kBlock * V];
  const bool vd = st->norm == 0;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int wi = xcd_block();
  const int slab = wi / ntiles, tile = wi - slab * ntiles;
  struct BC {
    VecN<double, V> B;
    VecN<float, V> C;
  };
  auto pre = [&](int P) { return BC{ldv<double, V>(b + P), ldv<float, V>(c32 + P)}; };
  auto epi = [&](int P, int i, int j, int k, const Gath<double, V> &G, const BC &pv) {
    const VecN<double, V> &Bv = pv.B;
    const VecN<double, V> R = resid<double, V>(Bv, G);
    const VecN<float, V> &Cf = pv.C;
    VecN<float, V> Rf, E;
#pragma unroll
    for (int u = 0; u < V; u++) {
      Rf.v[u] = (float)R.v[u];
      E.v[u] = omega * Rf.v[u] / Cf.v[u];
      acc[0] += R.v[u] * R.v[u];
      if (vd) {
        const double dv = m.dV(i, j, k + u);
        acc[1] += R.v[u] * R.v[u] / dv / dv;
      }
      acc[2] += (double)Rf.v[u] * (double)Rf.v[u];
      acc[3] += Bv.v[u] * Bv.v[u];
    }
    stv<float, V>(r32 + P, Rf);
    stv<float, V>(e1 + P, E);
  };
  lds_march_item<true, double, V, SYM, false, double>(g, A, ArrX<double, V>{x}, TI, slab, tile,
                                                      buf, pre, epi);
  block_partial<4>(acc, w.partial);
}

}  // namespace psb
