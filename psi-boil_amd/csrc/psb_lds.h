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

// one (slab, tile) item: stage, gather and epi(P, i, j, k, G) per plane.
// Every thread of the block runs the same number of planes (the LDS
// barriers); threads past the grid's last row only stage nothing.
template <typename T, int V, bool SYM, bool FLUX, typename M, class B, class Epi>
__device__ __forceinline__ void lds_march_item(const Grid &g, const DMat<M> &A, const B &base,
                                               int TI, int slab, int tile, T (*buf)[2 * kBlock * V],
                                               Epi epi) {
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
  const int i0 = slab * TI, i1 = min(i0 + TI, g.nx);
  int P = (i0 * g.ny + j) * g.nz + k;
  int HP = (i0 * g.ny + hj) * g.nz + hk;
  VecN<T, V> prev = zerov<T, V>(), cur = zerov<T, V>(), hcur = zerov<T, V>();
  if (act) {
    if (i0 > 0) prev = base.vec(P - g.sI, i0 - 1, j, k);
    cur = base.vec(P, i0, j, k);
  }
  if (hact) hcur = base.vec(HP, i0, hj, hk);
  for (int i = i0; i < i1; i++, P += g.sI, HP += g.sI) {
    VecN<T, V> next = zerov<T, V>(), hnext = zerov<T, V>();
    if (act && i + 1 < g.nx) next = base.vec(P + g.sI, i + 1, j, k);
    if (hact && i + 1 < g.nx) hnext = base.vec(HP + g.sI, i + 1, hj, hk);
    T *rows = buf[i & 1];
    if (act) *reinterpret_cast<VecN<T, V> *>(rows + (r + 1) * g.nz + k) = cur;
    if (hact) *reinterpret_cast<VecN<T, V> *>(rows + (hr + 1) * g.nz + hk) = hcur;
    __syncthreads();
    if (act) {
      const LdsX<T, V> X{rows, prev, cur, next, i, j, j0, g.nz};
      const Gath<T, V> G = gather<T, V, SYM, FLUX>(g, A, X, P, i, j, k);
      epi(P, i, j, k, G);
    }
    prev = cur;
    cur = next;
    hcur = hnext;
  }
}

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
  auto epi = [&](int P, int, int, int, const Gath<T, V> &G) {
    const VecN<T, V> R = ldv<T, V>(r + P);
    const VecN<T, V> t = resid<T, V>(R, G);
    VecN<T, V> E;
#pragma unroll
    for (int u = 0; u < V; u++) {
      E.v[u] = G.x.v[u] + omega * t.v[u] / G.c.v[u];
      if (RHO) {
        E.v[u] = zscale * E.v[u];
        rho_acc<T>(acc, R.v[u], E.v[u], qf, P + u, !af || af[P + u]);
      }
    }
    stv<T, V>(eout + P, E);
  };
  if constexpr (PROL)
    lds_march_item<T, V, true, FLUX, M>(g, A, pr, TI, slab, tile, buf, epi);
  else
    lds_march_item<T, V, true, FLUX, M>(g, A, ArrX<T, V>{e}, TI, slab, tile, buf, epi);
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
  auto epi = [&](int P, int, int, int, const Gath<T, V> &G) {
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
    lds_march_item<T, V, SYM, FLUX, M>(g, A, base, TI, slab, tile, buf, epi);
  } else {
    const ZPX<T, V> base{r_or_z, pold, af, (T)st->beta, (T)st->zmean};
    lds_march_item<T, V, SYM, FLUX, M>(g, A, base, TI, slab, tile, buf, epi);
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
  auto epi = [&](int P, int i, int j, int k, const Gath<double, V> &G) {
    const VecN<double, V> Bv = ldv<double, V>(b + P);
    const VecN<double, V> R = resid<double, V>(Bv, G);
    const VecN<float, V> Cf = ldv<float, V>(c32 + P);
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
  lds_march_item<double, V, SYM, false, double>(g, A, ArrX<double, V>{x}, TI, slab, tile, buf,
                                                epi);
  block_partial<4>(acc, w.partial);
}

}  // namespace psb
