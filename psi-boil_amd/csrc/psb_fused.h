// psb_fused.h -- two Jacobi sweeps in one pass over the fine level.
//
// The two post-smoothing sweeps of the V-cycle (the first with the coarse
// correction prolongated into its operand, the second producing the
// preconditioner output z and the partial sums for rho) read the matrix, r
// and e from HBM twice and write the intermediate field once when run as two
// kernels.  Here a block owns a tile of R rows (j) times the whole k range
// and marches through a slab of planes (i): the first sweep is computed
// plane by plane on the tile plus a one-row / one-plane halo, kept for the
// three planes the second sweep needs (the thread's own column in
// registers, the current plane in LDS for its j and k neighbours), and the
// second sweep is applied to the tile's interior.  Every value is computed
// with the same operations as by the two separate kernels, so z is
// bit-identical; only the block partition of the rho partial sums differs.
//
// Measured on MI355X (256^3, FP16 matrix): 550 us for the pass against
// 211 us for the two separate sweeps; 128^3: 100 against 24 us.  The march
// is latency-bound (one plane in flight per block, 183 VGPRs, two waves per
// SIMD, two barriers per plane) and the slab halo recomputes the first
// sweep on (TI + 2) / TI planes, so the solver keeps the two kernels unless
// the caller asks for this form (psb_stop.ac_fuse = 1).
#pragma once

#include "psb_kernels.h"

namespace psb {

#ifndef PSB_FUSE_BLOCKS
#define PSB_FUSE_BLOCKS 1536
#endif
#ifndef PSB_FUSE_MIN_TI
#define PSB_FUSE_MIN_TI 2
#endif
constexpr int kFuseBlocks = PSB_FUSE_BLOCKS;   // target grid of the fused pass
constexpr int kFuseMinTI = PSB_FUSE_MIN_TI;    // fewest planes per slab
constexpr int kFuseItems = 4;                  // cell vectors per thread and plane
constexpr int kFuseCells = kBlock * kFuseItems * 4;  // floats of one LDS plane

// the first sweep's values as the operand of the second: own column in
// registers (planes i-1, i, i+1), the current plane's j and k neighbours in
// LDS (row r of the tile at lds[r * nz + k])
struct RingX {
  const float *lds;
  VecN<float, 4> prev, cur, next;
  int i, j, j0, nz;  // j0: the tile's first row (LDS row 0 is j0 - 1)
  __device__ __forceinline__ VecN<float, 4> vec(int, int ii, int jj, int k) const {
    if (ii < i) return prev;
    if (ii > i) return next;
    if (jj == j) return cur;
    return ldv<float, 4>(lds + (jj - j0 + 1) * nz + k);
  }
  __device__ __forceinline__ float at(int, int, int jj, int k) const {
    return lds[(jj - j0 + 1) * nz + k];
  }
};
__device__ __forceinline__ VecN<float, 4> x_nbj(const RingX &x, int, int, int jj, int k) {
  return ldv<float, 4>(x.lds + (jj - x.j0 + 1) * x.nz + k);
}
__device__ __forceinline__ VecN<float, 4> x_nbi(const RingX &x, bool up, int, int, int, int) {
  return up ? x.next : x.prev;
}

// R: tile rows including the two halo rows (R * nz / 4 <= kBlock * kFuseItems);
// TI: planes per slab; ntiles: tiles along j.  Block b works on slab
// w / ntiles, tile w % ntiles with w the XCD-remapped block index, so an XCD
// holds neighbouring tiles of the same slabs (shared halo rows in its L2).
template <typename M, bool FLUX, bool RHO>
__global__ __launch_bounds__(kBlock) void k_post2(Grid g, DMat<M> A, const float *r,
                                                  float *z, ProlX<float, 4> pr, float omega,
                                                  const uint8_t *af, DevState *st, Work w,
                                                  float zscale, int R, int TI, int ntiles,
                                                  const float *qf) {
  if (st->done) return;
  __shared__ float lds[kFuseCells];
  const int nz = g.nz, nv = nz / 4;
  const int TJ = R - 2;
  const int wi = xcd_block();
  const int slab = wi / ntiles, tile = wi - slab * ntiles;
  const int i0 = slab * TI, i1 = min(i0 + TI, g.nx);
  const int j0 = tile * TJ;
  const int nitem = R * nv;
  // the first sweep on planes i-1, i, i+1 of the thread's items
  VecN<float, 4> prv[kFuseItems], cur[kFuseItems], nxt[kFuseItems];
  double acc[kRhoParts] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};

  // first sweep at plane ip for every item of the tile (zero outside the grid)
  auto sweep1 = [&](int ip, VecN<float, 4>(&out)[kFuseItems]) {
#pragma unroll
    for (int m = 0; m < kFuseItems; m++) {
      const int q = threadIdx.x + m * kBlock;
      VecN<float, 4> E = zerov<float, 4>();
      if (q < nitem && ip >= 0 && ip < g.nx) {
        const int row = q / nv, k = 4 * (q - row * nv);
        const int j = j0 - 1 + row;
        if (j >= 0 && j < g.ny) {
          const int P = (ip * g.ny + j) * nz + k;
          Gath<float, 4> G = gather<float, 4, true, FLUX>(g, A, pr, P, ip, j, k);
          VecN<float, 4> Rv = ldv<float, 4>(r + P);
          VecN<float, 4> t = resid<float, 4>(Rv, G);
#pragma unroll
          for (int u = 0; u < 4; u++) E.v[u] = G.x.v[u] + omega * t.v[u] / G.c.v[u];
        }
      }
      out[m] = E;
    }
  };
  auto to_lds = [&](const VecN<float, 4>(&v)[kFuseItems]) {
#pragma unroll
    for (int m = 0; m < kFuseItems; m++) {
      const int q = threadIdx.x + m * kBlock;
      if (q < nitem) {
        const int row = q / nv, k = 4 * (q - row * nv);
        stv<float, 4>(lds + row * nz + k, v[m]);
      }
    }
  };

  sweep1(i0 - 1, prv);
  sweep1(i0, cur);
  to_lds(cur);
  __syncthreads();
  for (int i = i0; i < i1; i++) {
    sweep1(i + 1, nxt);
    // second sweep on the interior rows of plane i
#pragma unroll
    for (int m = 0; m < kFuseItems; m++) {
      const int q = threadIdx.x + m * kBlock;
      if (q >= nitem) continue;
      const int row = q / nv, k = 4 * (q - row * nv);
      const int j = j0 - 1 + row;
      if (row < 1 || row > R - 2 || j >= g.ny) continue;
      const int P = (i * g.ny + j) * nz + k;
      RingX X{lds, prv[m], cur[m], nxt[m], i, j, j0, nz};
      Gath<float, 4> G = gather<float, 4, true, FLUX>(g, A, X, P, i, j, k);
      VecN<float, 4> Rv = ldv<float, 4>(r + P);
      VecN<float, 4> t = resid<float, 4>(Rv, G), E;
#pragma unroll
      for (int u = 0; u < 4; u++) {
        E.v[u] = G.x.v[u] + omega * t.v[u] / G.c.v[u];
        if (RHO) {
          E.v[u] = zscale * E.v[u];
          rho_acc<float>(acc, Rv.v[u], E.v[u], qf, P + u, !af || af[P + u]);
        }
      }
      stv<float, 4>(z + P, E);
    }
    __syncthreads();  // everyone is done with plane i in LDS
    to_lds(nxt);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kFuseItems; m++) {
      prv[m] = cur[m];
      cur[m] = nxt[m];
    }
  }
  if (RHO) block_partial<kRhoParts>(acc, w.partial);
}

// launch geometry of k_post2 for a level, or false when it does not apply
struct FuseShape {
  int R, TI, ntiles, nblocks;
};
inline bool fuse_shape(const Grid &g, FuseShape &f) {
  if (g.nz % 4 != 0 || g.N < 32768) return false;
  const int nv = g.nz / 4;
  int R = (kBlock * kFuseItems) / nv;
  if (R > g.ny + 2) R = g.ny + 2;
  if (R < 3) return false;
  f.R = R;
  f.ntiles = (g.ny + R - 3) / (R - 2);
  // about kFuseBlocks blocks, at most kMaxBlocks (the partials buffer)
  int TI = (int)((f.ntiles * (long long)g.nx + kFuseBlocks - 1) / kFuseBlocks);
  if (TI < kFuseMinTI) TI = kFuseMinTI;
  while (f.ntiles * ((g.nx + TI - 1) / TI) > kMaxBlocks) TI++;
  f.TI = TI;
  f.nblocks = f.ntiles * ((g.nx + TI - 1) / TI);
  return true;
}

}  // namespace psb
