// psb_api.hip -- host side of the MI355X pressure solver and its C ABI.
//
// One context holds the grid, the device copy of the matrix (FP64, and FP32
// for mixed precision), the additive-correction hierarchy, the Krylov work
// vectors, a HIP stream and the captured graphs of the iteration bodies.
// The loops of Linear::solve (cg.cpp, bicgs.cpp) run on the device: each
// graph launch performs `graph_chunk` iterations, every reduction is a
// streaming kernel plus a one-block finalize kernel that applies the scalar
// step and the stop test, the streaming kernels test the device-side stop
// flag, and the host only polls that flag between graph launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>


#include "psb_fused.h"
#include "psb_march.h"
#include "psb_lds.h"
#include "psb_internal.h"

// FP32 levels evaluate the stencil in flux form (set to 0 for the plain form)
#ifndef PSB_FLUX
#define PSB_FLUX 1
#endif

using namespace psb;

using namespace psb::host;



namespace psb {
namespace host {

DMat<double> mat64(const psb_ctx *c) {
  return DMat<double>{c->A64[0], c->A64[1], c->A64[2], c->A64[3],
                      c->A64[4], c->A64[5], c->A64[6]};
}
DMat<float> mat32(const psb_ctx *c) {
  return DMat<float>{c->A32[0], c->A32[1], c->A32[2], c->A32[3], nullptr, nullptr, nullptr};
}
Metric metric(const psb_ctx *c) { return Metric{c->dxc, c->dyc, c->dzc}; }

// 16-byte vectors: V = 2 doubles or 4 floats when nz allows, else 1
template <typename T>
constexpr int vmax() { return sizeof(T) == 8 ? 2 : 4; }
template <typename T>
bool vec_ok(const Grid &g) { return g.nz % vmax<T>() == 0; }
template <typename T>
int nblocks(const Grid &g) { return grid_blocks(vec_ok<T>(g) ? g.N / vmax<T>() : g.N); }

// stencil kernels on levels of at least kMarchMin cells march through the
// i planes (psb_march.h); smaller levels use the span kernels
constexpr bool kMarch = true;
// the first post-sweep (prolongation fused in) marches in FP64; in FP32 with
// the FP16 matrix the span kernel measured faster (28.0 against 28.5 ms per
// 256^3 solve; FP64: 56.4 marching against 59.0 ms)
template <typename T>
constexpr bool march_prol() { return std::is_same<T, double>::value; }
#ifndef PSB_MARCH_MIN
#define PSB_MARCH_MIN (1 << 15)
#endif
constexpr int kMarchMin = PSB_MARCH_MIN;
// fine-level kernels that run in the LDS-staged marching form (psb_lds.h),
// per precision (psb_ctx::lds32 / lds64).  Measured at 256^3 in the solve:
// FP32 first post-sweep 84.8 against 98.5 us (span), CG p/q step 109 against
// 131 us (register march); the other FP32 sweeps were slower staged (97 /
// 121 against 85 / 113 us), so they keep the register march.
enum { kLdsSweep = 1, kLdsProl = 2, kLdsLast = 4, kLdsCgPq = 8, kLdsResid = 16 };
template <typename T>
bool lds_use(const psb_ctx *C, int kind) {
  return ((sizeof(T) == 8 ? C->lds64 : C->lds32) & kind) != 0;
}

// calls f(std::integral_constant<int, V>) with the vector width for g
template <typename T, typename F>
void with_v(const Grid &g, F f) {
  if (vec_ok<T>(g))
    f(std::integral_constant<int, vmax<T>()>());
  else
    f(std::integral_constant<int, 1>());
}

template <int NR>
void finalize(psb_ctx *C, int op, int nblk, DevState *st, DevState *st2 = nullptr,
              double *out = nullptr) {
  hipLaunchKernelGGL((k_finalize<NR>), dim3(1), dim3(kBlock), 0, C->stream, op, nblk,
                     C->work.partial, st, st2, out);
}

// Per-kernel timing of an eager solve (psb_profile_solve_device): a kernel
// on the finest level launched inside a scope is bracketed by two HIP
// events on the context's stream; `per_cell` algorithmic bytes per cell.
// Does nothing in ordinary solves or on coarse levels.
struct ProbeScope {
  psb_ctx *C = nullptr;
  int kind = 0;
  hipEvent_t e0 = nullptr;
  ProbeScope(psb_ctx *c, int k, const Grid &g, double per_cell) {
    if (!c->probe.on || g.N != c->g.N) return;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventRecord(e0, c->stream));
    C = c;
    kind = k;
    C->probe.bytes[k] = per_cell * (double)g.N;
  }
  // closes the bracket (before a kernel that is not to be counted)
  void end() {
    if (!C) return;
    hipEvent_t e1 = nullptr;
    if (hipEventCreate(&e1) == hipSuccess) {
      (void)hipEventRecord(e1, C->stream);
      C->probe.rec.emplace_back(kind, e0, e1);
    }
    C = nullptr;
  }
  ~ProbeScope() { end(); }
  ProbeScope(const ProbeScope &) = delete;
  ProbeScope &operator=(const ProbeScope &) = delete;
};

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T, bool SYM>
void spmv(hipStream_t stream, const Grid &g, DMat<T> A, const T *x, T *y) {
  with_v<T>(g, [&](auto vt) {
    constexpr int V = decltype(vt)::value;
    MarchShape m;
    if (kMarch && V > 1 && g.N >= kMarchMin && march_ok<V>(g, m)) {
      hipLaunchKernelGGL((k_spmv_march<T, V, SYM>), dim3(m.nblocks), dim3(kBlock), 0, stream, g,
                         A, x, y, m.TI, m.ntiles);
      return;
    }
    hipLaunchKernelGGL((k_spmv<T, V, SYM>), dim3(nblocks<T>(g)), dim3(kBlock), 0, stream, g, A,
                       x, y);
  });
}

template <typename T, bool SYM>
void resid_k(hipStream_t stream, const Grid &g, DMat<T> A, const T *b, const T *x, T *y) {
  with_v<T>(g, [&](auto vt) {
    constexpr int V = decltype(vt)::value;
    hipLaunchKernelGGL((k_resid<T, V, SYM>), dim3(nblocks<T>(g)), dim3(kBlock), 0, stream, g, A,
                       b, x, y);
  });
}

template <typename T, bool SYM, int MODE>
void start(psb_ctx *C, const Grid &g, DMat<T> A, const T *b, const T *x, T *r, T *aux, T omega,
           DevState *st, const Metric *mt = nullptr) {
  const Metric m = mt ? *mt : metric(C);
  with_v<T>(g, [&](auto vt) {
    constexpr int V = decltype(vt)::value;
    hipLaunchKernelGGL((k_start<T, V, SYM, MODE>), dim3(nblocks<T>(g)), dim3(kBlock), 0,
                       C->stream, g, A, m, b, x, r, aux, omega, C->work);
  });
  finalize<3>(C, MODE == 0 ? FIN_START_DIAG : MODE == 1 ? FIN_START_AC : FIN_START_BICGS,
              nblocks<T>(g), st);
}

template <typename T, bool SYM, bool DIAG, bool FLUX = false>
void cg_pq(psb_ctx *C, const Grid &g, DMat<T> A, const T *rz, const T *pold, T *pnew, T *q,
           const uint8_t *af, DevState *st) {
  int nb = nblocks<T>(g);
  // the stored coefficient arrays, z (or r), p in; p, q out
  ProbeScope ps(C, PSB_K_CG_PQ, g, ((SYM ? 4 : 7) + 4) * sizeof(T) + (af ? 1 : 0));
  with_v<T>(g, [&](auto vt) {
    constexpr int V = decltype(vt)::value;
    MarchShape m;
    if (kMarch && V > 1 && g.N >= kMarchMin && march_ok<V>(g, m)) {
      nb = m.nblocks;
      if (lds_use<T>(C, kLdsCgPq) && lds_ok<V>(g))
        hipLaunchKernelGGL((k_cg_pq_lds<T, V, SYM, DIAG, FLUX>), dim3(m.nblocks), dim3(kBlock),
                           0, C->stream, g, A, rz, pold, pnew, q, af, st, C->work, m.TI,
                           m.ntiles, T(1));
      else
        hipLaunchKernelGGL((k_cg_pq_march<T, V, SYM, DIAG, FLUX>), dim3(m.nblocks),
                           dim3(kBlock), 0, C->stream, g, A, rz, pold, pnew, q, af, st,
                           C->work, m.TI, m.ntiles);
      return;
    }
    hipLaunchKernelGGL((k_cg_pq<T, V, SYM, DIAG, FLUX>), dim3(nblocks<T>(g)), dim3(kBlock), 0,
                       C->stream, g, A, rz, pold, pnew, q, af, st, C->work);
  });
  ps.end();
  finalize<1>(C, FIN_PQ, nb, st);
}

template <typename T, bool DIAG>
void cg_xr(psb_ctx *C, const Grid &g, const T *c, const T *p, const T *q, T *x, T *r, T *e1,
           T omega, DevState *st, int op = -1, const Metric *mt = nullptr) {
  const Metric m = mt ? *mt : metric(C);
  {
    // x, r, p, q, c in; x, r (and e1) out
    ProbeScope ps(C, PSB_K_CG_XR, g, (DIAG ? 7 : 8) * sizeof(T));
    with_v<T>(g, [&](auto vt) {
      constexpr int V = decltype(vt)::value;
      hipLaunchKernelGGL((k_cg_xr<T, V, DIAG>), dim3(nblocks<T>(g)), dim3(kBlock), 0, C->stream,
                         g, c, m, p, q, x, r, e1, omega, st, C->work);
    });
  }
  finalize<3>(C, op >= 0 ? op : DIAG ? FIN_XR_DIAG : FIN_XR_AC, nblocks<T>(g), st);
}

// one damped Jacobi sweep on level lv (optionally with the prolongation of
// the coarse correction eH fused in; RHO: last post-sweep on level 0)
// one damped Jacobi sweep on level lv (optionally with the prolongation of
// the coarse correction eH fused in; RHO: last post-sweep on level 0, its
// output multiplied by zscale).  With a half-precision copy of the level
// matrix (lv.c16, scaled by lv.scale) the sweep reads that copy.
template <typename T, bool RHO, typename M, bool FLUX>
void jacobi_m(psb_ctx *C, const Level<T> &lv, DMat<M> A, const T *e, T *eout,
              const Level<T> *coarse, const T *eH, T cfac, T omega, T zscale, DevState *st,
              const T *qf) {
  const Grid &g = lv.g;
  int nb = nblocks<T>(g);
  // four coefficient arrays, r, e in, e out; the coarse correction (one
  // coarse value per block of cells), q (flexible beta), the active flags
  const double per_cell =
      4.0 * sizeof(M) + 3.0 * sizeof(T) +
      (coarse ? (double)sizeof(T) / (coarse->CI * coarse->CJ * coarse->CK) : 0.0) +
      (RHO && qf ? sizeof(T) : 0) + (lv.aflag && (coarse || RHO) ? 1 : 0);
  ProbeScope ps(C, coarse ? PSB_K_SWEEP_PROL : RHO ? PSB_K_SWEEP_LAST : PSB_K_SWEEP, g, per_cell);
  with_v<T>(g, [&](auto vt) {
    constexpr int V = decltype(vt)::value;
    ProlX<T, V> pr{};
    if (coarse) {
      pr.e = e;
      pr.eH = eH;
      pr.af = lv.aflag;
      pr.cfac = cfac;
      pr.sI = coarse->CI == 2 ? 1 : 0;
      pr.sJ = coarse->CJ == 2 ? 1 : 0;
      pr.sK = coarse->CK == 2 ? 1 : 0;
      pr.nYH = coarse->g.ny;
      pr.nZH = coarse->g.nz;
    }
    if constexpr (std::is_same<T, double>::value && std::is_same<M, T>::value && !FLUX &&
                  !RHO) {
      if (lv.ae) {  // non-symmetric level (7 arrays): the span kernel
        if (coarse)
          hipLaunchKernelGGL((k_jacobi<T, V, false, true, false, M, false>),
                             dim3(nblocks<T>(g)), dim3(kBlock), 0, C->stream, g, A, lv.r, e,
                             eout, pr, omega, lv.aflag, st, C->work, zscale, qf);
        else
          hipLaunchKernelGGL((k_jacobi<T, V, false, false, false, M, false>),
                             dim3(nblocks<T>(g)), dim3(kBlock), 0, C->stream, g, A, lv.r, e,
                             eout, pr, omega, lv.aflag, st, C->work, zscale, qf);
        return;
      }
    }
    MarchShape m;
    if (lds_use<T>(C, coarse ? kLdsProl : RHO ? kLdsLast : kLdsSweep) && V > 1 &&
        g.N >= kMarchMin && lds_ok<V>(g) && march_ok<V>(g, m)) {
      nb = m.nblocks;
      if (coarse)
        hipLaunchKernelGGL((k_jacobi_lds<T, V, true, RHO, M, FLUX>), dim3(m.nblocks),
                           dim3(kBlock), 0, C->stream, g, A, lv.r, e, eout, pr, omega, lv.aflag,
                           st, C->work, zscale, m.TI, m.ntiles, qf);
      else
        hipLaunchKernelGGL((k_jacobi_lds<T, V, false, RHO, M, FLUX>), dim3(m.nblocks),
                           dim3(kBlock), 0, C->stream, g, A, lv.r, e, eout, pr, omega, lv.aflag,
                           st, C->work, zscale, m.TI, m.ntiles, qf);
      return;
    }
    if (kMarch && V > 1 && g.N >= kMarchMin && (!coarse || march_prol<T>()) && march_ok<V>(g, m)) {
      nb = m.nblocks;
      if (coarse)
        hipLaunchKernelGGL((k_jacobi_march<T, V, true, RHO, M, FLUX>), dim3(m.nblocks),
                           dim3(kBlock), 0, C->stream, g, A, lv.r, e, eout, pr, omega, lv.aflag,
                           st, C->work, zscale, m.TI, m.ntiles, qf);
      else
        hipLaunchKernelGGL((k_jacobi_march<T, V, false, RHO, M, FLUX>), dim3(m.nblocks),
                           dim3(kBlock), 0, C->stream, g, A, lv.r, e, eout, pr, omega, lv.aflag,
                           st, C->work, zscale, m.TI, m.ntiles, qf);
      return;
    }
    if (coarse) {
      hipLaunchKernelGGL((k_jacobi<T, V, true, true, RHO, M, FLUX>), dim3(nblocks<T>(g)),
                         dim3(kBlock), 0, C->stream, g, A, lv.r, e, eout, pr, omega, lv.aflag,
                         st, C->work, zscale, qf);
    } else {
      hipLaunchKernelGGL((k_jacobi<T, V, true, false, RHO, M, FLUX>), dim3(nblocks<T>(g)),
                         dim3(kBlock), 0, C->stream, g, A, lv.r, e, eout, pr, omega, lv.aflag,
                         st, C->work, zscale, qf);
    }
  });
  ps.end();
  if (RHO) finalize<kRhoParts>(C, qf ? FIN_RHO_FLEX : FIN_RHO, nb, st);
}

// qf (RHO only): q = A p of the last search direction for the flexible beta
template <typename T, bool RHO>
void jacobi(psb_ctx *C, const Level<T> &lv, const T *e, T *eout, const Level<T> *coarse,
            const T *eH, T cfac, T omega, DevState *st, T zscale = T(1),
            const T *qf = nullptr) {
  if constexpr (std::is_same<T, float>::value) {
    if (lv.h16() && lv.d) {
      jacobi_m<T, RHO, _Float16, true>(C, lv, lv.mat16(), e, eout, coarse, eH, cfac, omega,
                                       zscale, st, qf);
      return;
    }
    if (lv.h16()) {
      jacobi_m<T, RHO, _Float16, false>(C, lv, lv.mat16(), e, eout, coarse, eH, cfac, omega,
                                        zscale, st, qf);
      return;
    }
    if (lv.d) {
      jacobi_m<T, RHO, T, true>(C, lv, lv.matflux(), e, eout, coarse, eH, cfac, omega, zscale,
                                st, qf);
      return;
    }
  }
  jacobi_m<T, RHO, T, false>(C, lv, lv.mat(), e, eout, coarse, eH, cfac, omega, zscale, st, qf);
}

// the two post-sweeps of level lv in one pass (k_post2): the first with the
// prolongation of eH from `coarse`, the second writing eout (and the rho
// partials when RHO).  False when the level's shape does not suit it.
template <bool RHO, typename M, bool FLUX>
void post2_m(psb_ctx *C, const Level<float> &lv, DMat<M> A, const float *e, float *eout,
             const Level<float> &coarse, const float *eH, float cfac, float omega, float zscale,
             DevState *st, const FuseShape &f, const float *qf) {
  ProlX<float, 4> pr{};
  pr.e = e;
  pr.eH = eH;
  pr.af = lv.aflag;
  pr.cfac = cfac;
  pr.sI = coarse.CI == 2 ? 1 : 0;
  pr.sJ = coarse.CJ == 2 ? 1 : 0;
  pr.sK = coarse.CK == 2 ? 1 : 0;
  pr.nYH = coarse.g.ny;
  pr.nZH = coarse.g.nz;
  hipLaunchKernelGGL((k_post2<M, FLUX, RHO>), dim3(f.nblocks), dim3(kBlock), 0, C->stream, lv.g,
                     A, lv.r, eout, pr, omega, lv.aflag, st, C->work, zscale, f.R, f.TI,
                     f.ntiles, qf);
  if (RHO) finalize<kRhoParts>(C, qf ? FIN_RHO_FLEX : FIN_RHO, f.nblocks, st);
}

template <typename T, bool RHO>
bool post2(psb_ctx *C, const Level<T> &lv, const T *e, T *eout, const Level<T> &coarse,
           const T *eH, T cfac, T omega, T zscale, DevState *st, const T *qf = nullptr) {
  if constexpr (std::is_same<T, float>::value) {
    FuseShape f;
    if (!C->fuse || !fuse_shape(lv.g, f)) return false;
    if (lv.h16() && lv.d)
      post2_m<RHO, _Float16, true>(C, lv, lv.mat16(), e, eout, coarse, eH, cfac, omega, zscale,
                                   st, f, qf);
    else if (lv.h16())
      post2_m<RHO, _Float16, false>(C, lv, lv.mat16(), e, eout, coarse, eH, cfac, omega, zscale,
                                    st, f, qf);
    else if (lv.d)
      post2_m<RHO, float, true>(C, lv, lv.matflux(), e, eout, coarse, eH, cfac, omega, zscale,
                                st, f, qf);
    else
      post2_m<RHO, float, false>(C, lv, lv.mat(), e, eout, coarse, eH, cfac, omega, zscale, st,
                                 f, qf);
    return true;
  } else {
    return false;
  }
}

template <typename T, int V, int KV, int CI, int CJ, int CK, typename M, bool FLUX>
void restrict_launch(psb_ctx *C, const Level<T> &f, DMat<M> A, const T *e, Level<T> &cL,
                     T omega, DevState *st) {
  // fine: four coefficient arrays, r, e (and flags); coarse: c in, r and e out
  ProbeScope ps(C, PSB_K_RESTRICT, f.g,
                4.0 * sizeof(M) + 2.0 * sizeof(T) + (f.aflag ? 1 : 0) +
                    (3.0 * sizeof(T) + (cL.aflag ? 1 : 0)) / (CI * CJ * CK));
  if constexpr (std::is_same<T, double>::value && std::is_same<M, T>::value && !FLUX) {
    if (f.ae) {  // non-symmetric level (7 arrays): the span kernel
      const int nwork = cL.g.nx * cL.g.ny * (f.g.nz / (V * KV));
      hipLaunchKernelGGL((k_restrict<T, V, KV, false, CI, CJ, CK, M, false>),
                         dim3(grid_blocks(nwork)), dim3(kBlock), 0, C->stream, f.g, A, f.r, e,
                         f.aflag, cL.g, cL.c, cL.aflag, cL.r, cL.e[0], omega, st);
      return;
    }
  }
  if constexpr (CI == 2 && KV == 1 && V > 1) {
    if (kMarch && f.g.N >= kMarchMin && f.g.nx == 2 * cL.g.nx) {
      const MarchShape m = march_shape_n(cL.g.nx, (long long)cL.g.ny * (f.g.nz / V));
      if (m.nblocks <= kMaxBlocks) {
        hipLaunchKernelGGL((k_restrict_march<T, V, true, CJ, CK, M, FLUX>), dim3(m.nblocks),
                           dim3(kBlock), 0, C->stream, f.g, A, f.r, e, f.aflag, cL.g, cL.c,
                           cL.aflag, cL.r, cL.e[0], omega, st, m.TI, m.ntiles);
        return;
      }
    }
  }
  const int nwork = cL.g.nx * cL.g.ny * (f.g.nz / (V * KV));
  hipLaunchKernelGGL((k_restrict<T, V, KV, true, CI, CJ, CK, M, FLUX>), dim3(grid_blocks(nwork)),
                     dim3(kBlock), 0, C->stream, f.g, A, f.r, e, f.aflag, cL.g, cL.c, cL.aflag,
                     cL.r, cL.e[0], omega, st);
}

template <typename T, int V, int KV, int CI, int CJ, int CK>
void restrict_v(psb_ctx *C, const Level<T> &f, const T *e, Level<T> &cL, T omega,
                DevState *st) {
  if constexpr (std::is_same<T, float>::value) {
    if (f.h16() && f.d) {
      restrict_launch<T, V, KV, CI, CJ, CK, _Float16, true>(C, f, f.mat16(), e, cL, omega, st);
      return;
    }
    if (f.h16()) {
      restrict_launch<T, V, KV, CI, CJ, CK, _Float16, false>(C, f, f.mat16(), e, cL, omega, st);
      return;
    }
    if (f.d) {
      restrict_launch<T, V, KV, CI, CJ, CK, T, true>(C, f, f.matflux(), e, cL, omega, st);
      return;
    }
  }
  restrict_launch<T, V, KV, CI, CJ, CK, T, false>(C, f, f.mat(), e, cL, omega, st);
}

template <typename T, int V, int CI, int CJ>
void restrict_k(psb_ctx *C, const Level<T> &f, const T *e, Level<T> &cL, T omega,
                DevState *st) {
  if constexpr (V == 1) {
    if (cL.CK == 2) restrict_v<T, 1, 2, CI, CJ, 2>(C, f, e, cL, omega, st);
    else restrict_v<T, 1, 1, CI, CJ, 1>(C, f, e, cL, omega, st);
  } else {
    if (cL.CK == 2) restrict_v<T, V, 1, CI, CJ, 2>(C, f, e, cL, omega, st);
    else restrict_v<T, V, 1, CI, CJ, 1>(C, f, e, cL, omega, st);
  }
}

// restriction of (r - A e) from level f to cL, with cL's first sweep fused
template <typename T>
void restrict_(psb_ctx *C, const Level<T> &f, const T *e, Level<T> &cL, T omega, DevState *st) {
  with_v<T>(f.g, [&](auto vt) {
    constexpr int V = decltype(vt)::value;
    if (cL.CI == 2 && cL.CJ == 2) restrict_k<T, V, 2, 2>(C, f, e, cL, omega, st);
    else if (cL.CI == 2) restrict_k<T, V, 2, 1>(C, f, e, cL, omega, st);
    else if (cL.CJ == 2) restrict_k<T, V, 1, 2>(C, f, e, cL, omega, st);
    else restrict_k<T, V, 1, 1>(C, f, e, cL, omega, st);
  });
}

// ---------------------------------------------------------------------------
// dense coarsest operator: G = Pi (A_pinned)^-1 Pi over active cells.  A
// Cholesky factorisation that fixes to zero every unknown whose pivot
// vanishes (the constant null direction of each connected Neumann block),
// then the inverse of the factor pair, then projection onto zero mean (the
// projection also removes what a last pivot that rounding left just above
// the pinning threshold put along the constant); project false (a
// non-singular operator: Dirichlet sides) keeps the inverse.
// ---------------------------------------------------------------------------
std::vector<double> dense_pinv(bool project, int nx, int ny, int nz, const std::vector<double> &c,
                               const std::vector<double> &w, const std::vector<double> &s,
                               const std::vector<double> &b, const std::vector<uint8_t> &act) {
  const int m = nx * ny * nz;
  std::vector<double> M((size_t)m * m, 0.0);
  auto id = [&](int i, int j, int k) { return (i * ny + j) * nz + k; };
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        int P = id(i, j, k);
        if (!act[P]) continue;
        M[(size_t)P * m + P] = c[P];
        auto cpl = [&](int Q, double a) {
          if (act[Q]) M[(size_t)P * m + Q] = -a;
        };
        if (i > 0) cpl(id(i - 1, j, k), w[P]);
        if (i < nx - 1) cpl(id(i + 1, j, k), w[id(i + 1, j, k)]);
        if (j > 0) cpl(id(i, j - 1, k), s[P]);
        if (j < ny - 1) cpl(id(i, j + 1, k), s[id(i, j + 1, k)]);
        if (k > 0) cpl(id(i, j, k - 1), b[P]);
        if (k < nz - 1) cpl(id(i, j, k + 1), b[id(i, j, k + 1)]);
      }
  std::vector<uint8_t> pin(m, 0);
  for (int P = 0; P < m; P++) pin[P] = act[P] ? 0 : 1;
  // the operator couples P to P +- 1, +- nz, +- ny nz: its Cholesky factor
  // stays inside that band (bw below the diagonal), so the loops skip the
  // entries outside it, which are exact zeros (same sums, same rounding)
  const int bw = nx > 1 ? ny * nz : (ny > 1 ? nz : 1);
  // row-by-row Cholesky (lower triangle in place, rows contiguous); an
  // unknown whose pivot vanishes is pinned: its row and column drop out
  for (int i = 0; i < m; i++) {
    double *Li = &M[(size_t)i * m];
    const int lo_i = std::max(0, i - bw);
    if (pin[i]) {
      for (int j = 0; j <= i; j++) Li[j] = 0.0;
      continue;
    }
    for (int j = lo_i; j <= i; j++) {
      if (pin[j]) {
        Li[j] = 0.0;
        continue;
      }
      const double *Lj = &M[(size_t)j * m];
      double sm = Li[j];
      for (int k2 = std::max(lo_i, j - bw); k2 < j; k2++) sm -= Li[k2] * Lj[k2];
      if (j < i) {
        Li[j] = sm / Lj[j];
      } else if (sm <= 1e-10 * std::fabs(c[i])) {
        pin[i] = 1;
        for (int q = 0; q <= i; q++) Li[q] = 0.0;
      } else {
        Li[i] = std::sqrt(sm);
      }
    }
  }
  // columns of the inverse: two banded triangular solves each (the forward
  // one starts at the column's own row: y is zero above it), independent,
  // on the host's threads (the coarsest level of a stretched grid can hold
  // 12^3 = 1728 cells)
  std::vector<double> B((size_t)m * m, 0.0);
  auto columns = [&](int c0, int c1) {
    std::vector<double> y(m);
    for (int col = c0; col < c1; col++) {
      if (pin[col]) continue;
      for (int i = 0; i < m; i++) y[i] = (i == col) ? 1.0 : 0.0;
      for (int i = col; i < m; i++) {
        if (pin[i]) { y[i] = 0.0; continue; }
        double sm = y[i];
        for (int k2 = std::max(col, i - bw); k2 < i; k2++)
          if (!pin[k2]) sm -= M[(size_t)i * m + k2] * y[k2];
        y[i] = sm / M[(size_t)i * m + i];
      }
      // L^T y = y by columns of L^T (rows of L, contiguous)
      for (int i = m - 1; i >= 0; i--) {
        if (pin[i]) { y[i] = 0.0; continue; }
        const double *Li = &M[(size_t)i * m];
        y[i] /= Li[i];
        const double yi = y[i];
        for (int k2 = std::max(0, i - bw); k2 < i; k2++) y[k2] -= Li[k2] * yi;
      }
      for (int i = 0; i < m; i++) B[(size_t)i * m + col] = y[i];
    }
  };
  const int nt = m < 256 ? 1 : (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; t++)
    pool.emplace_back(columns, (int)((long long)m * t / nt), (int)((long long)m * (t + 1) / nt));
  for (auto &th : pool) th.join();
  if (!project) return B;  // a non-singular operator: the inverse itself
  int na = 0;
  for (int P = 0; P < m; P++) na += act[P] ? 1 : 0;
  std::vector<double> rm(m, 0.0), cm(m, 0.0);
  double tm = 0.0;
  for (int i = 0; i < m; i++) {
    if (!act[i]) continue;
    for (int j = 0; j < m; j++) {
      if (!act[j]) continue;
      rm[i] += B[(size_t)i * m + j];
      cm[j] += B[(size_t)i * m + j];
      tm += B[(size_t)i * m + j];
    }
  }
  for (int i = 0; i < m; i++) {
    rm[i] /= na;
    cm[i] /= na;
  }
  tm /= (double)na * (double)na;
  std::vector<double> G((size_t)m * m, 0.0);
  for (int i = 0; i < m; i++)
    for (int j = 0; j < m; j++)
      if (act[i] && act[j]) G[(size_t)i * m + j] = B[(size_t)i * m + j] - rm[i] - cm[j] + tm;
  return G;
}

// level sizes: halve a direction while it stays >= min_dim (8: the
// benchmark's rule down to 8^3).  rule 1 halves x and y first and z only once
// x and y stop (semi-coarsening for grids stretched in z, where the
// couplings in x and y dominate away from the walls).  rule 2 is the
// reference's own coarsening (used by its AC cycle).
int smallest_factor(int n) {
  for (int f = 2; f * f <= n; f++)
    if (n % f == 0) return f;
  return n;
}

std::vector<int> level_dims(int nx, int ny, int nz, int min_dim, int rule) {
  std::vector<int> d = {nx, ny, nz};
  while ((int)d.size() / 3 < kMaxLevels) {
    int n[3] = {d[d.size() - 3], d[d.size() - 2], d[d.size() - 1]};
    int m[3];
    bool any = false;
    for (int q = 0; q < 3; q++) {
      m[q] = n[q];
      if (rule == 2) {
        // the reference's rule (domain_coarsen.cpp:33-71 with factor(),
        // global_factor.cpp:2-19): divide by the smallest prime factor while
        // the size exceeds 4 and is not itself that factor
        const int f = smallest_factor(n[q]);
        if (n[q] > 4 && n[q] != f) {
          m[q] = n[q] / f;
          any = true;
        }
        continue;
      }
      if (rule == 1 && q == 2 && any) continue;
      if (n[q] % 2 == 0 && n[q] / 2 >= min_dim) {
        m[q] = n[q] / 2;
        any = true;
      }
    }
    if (!any) break;
    d.push_back(m[0]);
    d.push_back(m[1]);
    d.push_back(m[2]);
  }
  return d;
}

template <typename T>
void ensure_krylov(psb_ctx *C, KrylovVecs<T> &K, bool bicgs) {
  const int N = C->g.N;
  if (!K.x) {
    K.x = C->arena.alloc<T>(N);
    K.r = C->arena.alloc<T>(N);
    K.q = C->arena.alloc<T>(N);
    K.p[0] = C->arena.alloc<T>(N);
    K.p[1] = C->arena.alloc<T>(N);
  }
  if (bicgs && !K.s) {
    K.v[0] = C->arena.alloc<T>(N);
    K.v[1] = C->arena.alloc<T>(N);
    K.s = C->arena.alloc<T>(N);
    K.t = C->arena.alloc<T>(N);
    K.rt = C->arena.alloc<T>(N);
  }
}

template void ensure_krylov<float>(psb_ctx *, KrylovVecs<float> &, bool);
template void ensure_krylov<double>(psb_ctx *, KrylovVecs<double> &, bool);

// dense coarsest operator of a non-symmetric matrix (7 arrays): G = A^-1 on
// the active cells (0 elsewhere) by LU with partial pivoting; such a level
// carries a time term, and a singular one is an error
std::vector<double> dense_inv(int nx, int ny, int nz, const std::vector<double> &c,
                              const std::vector<double> &w, const std::vector<double> &s,
                              const std::vector<double> &b, const std::vector<double> &e,
                              const std::vector<double> &n, const std::vector<double> &t,
                              const std::vector<uint8_t> &act) {
  const int m = nx * ny * nz;
  auto id = [&](int i, int j, int k) { return (i * ny + j) * nz + k; };
  std::vector<int> idx, pos(m, -1);  // active cells, in order
  for (int P = 0; P < m; P++)
    if (act[P]) { pos[P] = (int)idx.size(); idx.push_back(P); }
  const int na = (int)idx.size();
  std::vector<double> U((size_t)na * na, 0.0);
  double cmax = 0.0;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        const int P = id(i, j, k);
        if (!act[P]) continue;
        double *row = &U[(size_t)pos[P] * na];
        row[pos[P]] = c[P];
        cmax = std::max(cmax, std::fabs(c[P]));
        auto cpl = [&](int Q, double a) {
          if (act[Q]) row[pos[Q]] = -a;
        };
        if (i > 0) cpl(id(i - 1, j, k), w[P]);
        if (i < nx - 1) cpl(id(i + 1, j, k), e[P]);
        if (j > 0) cpl(id(i, j - 1, k), s[P]);
        if (j < ny - 1) cpl(id(i, j + 1, k), n[P]);
        if (k > 0) cpl(id(i, j, k - 1), b[P]);
        if (k < nz - 1) cpl(id(i, j, k + 1), t[P]);
      }
  std::vector<int> perm(na);
  for (int q = 0; q < na; q++) perm[q] = q;
  for (int k = 0; k < na; k++) {
    int p = k;
    for (int i = k + 1; i < na; i++)
      if (std::fabs(U[(size_t)i * na + k]) > std::fabs(U[(size_t)p * na + k])) p = i;
    if (!(std::fabs(U[(size_t)p * na + k]) > 1e-14 * cmax))
      throw std::invalid_argument("singular coarsest level of a non-symmetric matrix");
    if (p != k) {
      for (int j = 0; j < na; j++) std::swap(U[(size_t)p * na + j], U[(size_t)k * na + j]);
      std::swap(perm[p], perm[k]);
    }
    const double piv = U[(size_t)k * na + k];
    for (int i = k + 1; i < na; i++) {
      double *Ui = &U[(size_t)i * na];
      const double l = Ui[k] / piv;
      Ui[k] = l;
      if (l == 0.0) continue;
      const double *Uk = &U[(size_t)k * na];
      for (int j = k + 1; j < na; j++) Ui[j] -= l * Uk[j];
    }
  }
  std::vector<double> G((size_t)m * m, 0.0), y(na);
  for (int col = 0; col < na; col++) {
    for (int i = 0; i < na; i++) y[i] = perm[i] == col ? 1.0 : 0.0;
    for (int i = 0; i < na; i++) {
      double v = y[i];
      for (int k = 0; k < i; k++) v -= U[(size_t)i * na + k] * y[k];
      y[i] = v;
    }
    for (int i = na - 1; i >= 0; i--) {
      double v = y[i];
      for (int k = i + 1; k < na; k++) v -= U[(size_t)i * na + k] * y[k];
      y[i] = v / U[(size_t)i * na + i];
    }
    for (int i = 0; i < na; i++) G[(size_t)idx[i] * m + idx[col]] = y[i];
  }
  return G;
}

void build_hier64(psb_ctx *C) {
  Hier<double> &H = C->h64;
  if (H.built) return;
  ensure_krylov(C, C->k64, false);
  hipStream_t stream = C->stream;
  std::vector<int> d = level_dims(C->g.nx, C->g.ny, C->g.nz, C->coarse_min, C->coarse_rule);
  const int nl = (int)d.size() / 3;
  if (nl < 2) throw std::invalid_argument("grid too small for the multigrid preconditioner");
  H.L.resize(nl);
  for (int l = 0; l < nl; l++) {
    Level<double> &lv = H.L[l];
    lv.g = make_grid(d[3 * l], d[3 * l + 1], d[3 * l + 2]);
    if (l > 0) {
      lv.CI = d[3 * (l - 1)] / lv.g.nx;
      lv.CJ = d[3 * (l - 1) + 1] / lv.g.ny;
      lv.CK = d[3 * (l - 1) + 2] / lv.g.nz;
    }
    const size_t N = lv.g.N;
    if (l == 0) {
      lv.c = C->A64[0]; lv.w = C->A64[1]; lv.s = C->A64[2]; lv.b = C->A64[3];
      if (!C->sym) { lv.ae = C->A64[4]; lv.an = C->A64[5]; lv.at = C->A64[6]; }
      lv.r = C->k64.r;
      lv.aflag = C->aflag0;
    } else {
      lv.c = C->harena.alloc<double>(N); lv.w = C->harena.alloc<double>(N);
      lv.s = C->harena.alloc<double>(N); lv.b = C->harena.alloc<double>(N);
      if (!C->sym) {  // a non-symmetric matrix keeps 7 arrays on every level
        lv.ae = C->harena.alloc<double>(N); lv.an = C->harena.alloc<double>(N);
        lv.at = C->harena.alloc<double>(N);
      }
      lv.r = C->harena.alloc<double>(N);
      Level<double> &f = H.L[l - 1];
      hipLaunchKernelGGL(k_galerkin, dim3(grid_blocks(N)), dim3(kBlock), 0, stream, f.g,
                         f.mat(), lv.g, lv.CI, lv.CJ, lv.CK, lv.c, lv.w, lv.s, lv.b, lv.ae,
                         lv.an, lv.at);
      if (f.aflag) {
        lv.aflag = C->harena.alloc<uint8_t>(N);
        hipLaunchKernelGGL(k_coarsen_flag, dim3(grid_blocks(N)), dim3(kBlock), 0, stream, f.g,
                           f.aflag, lv.g, lv.CI, lv.CJ, lv.CK, lv.aflag);
      }
    }
    for (int q = 0; q < 3; q++) lv.e[q] = C->harena.alloc<double>(N);
  }
  HIPCHK(hipGetLastError());
  // coarsest dense operator
  Level<double> &cl = H.L[nl - 1];
  const int m = cl.g.N;
  std::vector<double> c(m), w(m), s(m), b(m);
  std::vector<uint8_t> act(m, 1);
  HIPCHK(hipMemcpyAsync(c.data(), cl.c, m * 8, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipMemcpyAsync(w.data(), cl.w, m * 8, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipMemcpyAsync(s.data(), cl.s, m * 8, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipMemcpyAsync(b.data(), cl.b, m * 8, hipMemcpyDeviceToHost, stream));
  if (cl.aflag) HIPCHK(hipMemcpyAsync(act.data(), cl.aflag, m, hipMemcpyDeviceToHost, stream));
  std::vector<double> e, n, t;
  if (cl.ae) {
    e.resize(m); n.resize(m); t.resize(m);
    HIPCHK(hipMemcpyAsync(e.data(), cl.ae, m * 8, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(n.data(), cl.an, m * 8, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(t.data(), cl.at, m * 8, hipMemcpyDeviceToHost, stream));
  }
  HIPCHK(hipStreamSynchronize(stream));
  std::vector<double> G = cl.ae ? dense_inv(cl.g.nx, cl.g.ny, cl.g.nz, c, w, s, b, e, n, t, act)
                                : dense_pinv(C->singular, cl.g.nx, cl.g.ny, cl.g.nz, c, w, s, b,
                                             act);
  H.m = m;
  H.G = C->harena.alloc<double>((size_t)m * m);
  HIPCHK(hipMemcpyAsync(H.G, G.data(), (size_t)m * m * 8, hipMemcpyHostToDevice, stream));
  HIPCHK(hipStreamSynchronize(stream));
  H.built = true;
}

void ensure32(psb_ctx *C) {
  if (C->have32) return;
  const int N = C->g.N;
  for (int q = 0; q < 4; q++) {
    C->A32[q] = C->arena.alloc<float>(N);
    hipLaunchKernelGGL(k_to_f32, dim3(grid_blocks(N)), dim3(kBlock), 0, C->stream, C->A64[q],
                       C->A32[q], N);
  }
  HIPCHK(hipGetLastError());
  C->have32 = true;
}

void build_hier32(psb_ctx *C) {
  Hier<float> &H = C->h32;
  if (H.built) return;
  build_hier64(C);
  ensure32(C);
  ensure_krylov(C, C->k32, false);
  const Hier<double> &H64 = C->h64;
  const int nl = (int)H64.L.size();
  H.L.resize(nl);
  for (int l = 0; l < nl; l++) {
    const Level<double> &s = H64.L[l];
    Level<float> &lv = H.L[l];
    lv.g = s.g;
    lv.CI = s.CI; lv.CJ = s.CJ; lv.CK = s.CK;
    lv.aflag = s.aflag;
    const int N = lv.g.N;
    if (l == 0) {
      lv.c = C->A32[0]; lv.w = C->A32[1]; lv.s = C->A32[2]; lv.b = C->A32[3];
      lv.r = C->k32.r;
    } else {
      float **dst[4] = {&lv.c, &lv.w, &lv.s, &lv.b};
      const double *src[4] = {s.c, s.w, s.s, s.b};
      for (int q = 0; q < 4; q++) {
        *dst[q] = C->harena.alloc<float>(N);
        hipLaunchKernelGGL(k_to_f32, dim3(grid_blocks(N)), dim3(kBlock), 0, C->stream, src[q],
                           *dst[q], N);
      }
      lv.r = C->harena.alloc<float>(N);
    }
    for (int q = 0; q < 3; q++) lv.e[q] = C->harena.alloc<float>(N);
#if PSB_FLUX
    // flux-form diagonal for the FP32 stencils
    lv.d = C->harena.alloc<float>(N);
    hipLaunchKernelGGL(k_flux_d<float>, dim3(grid_blocks(N)), dim3(kBlock), 0, C->stream, lv.g,
                       s.mat(), lv.d, 1.0);
#endif
  }
  H.G = H64.G;
  H.m = H64.m;
  HIPCHK(hipGetLastError());
  H.built = true;
}

// Whether the FP32 levels' smoother reads half-precision copies
// scale16 * (c, w, s, b) (psb_stop.ac_store >= 1): built on every level but
// the coarsest the first time a solve asks for them, then kept beside the
// FP32 matrices, so that solves with and without them alternate on one
// hierarchy (and their captured graphs stay valid).
void set_half(psb_ctx *C, bool on) {
  Hier<float> &H = C->h32;
  const Hier<double> &H64 = C->h64;
  const int nl = (int)H.L.size();
  for (int l = 0; l < nl - 1; l++) {
    Level<float> &lv = H.L[l];
    const Level<double> &s = H64.L[l];
    const int N = lv.g.N;
    if (on && !lv.c16) {
      const int nb = grid_blocks(N);
      hipLaunchKernelGGL(k_blockmax, dim3(nb), dim3(kBlock), 0, C->stream, s.c, N,
                         C->work.partial);
      std::vector<double> bm(nb);
      HIPCHK(hipMemcpyAsync(bm.data(), C->work.partial, nb * 8, hipMemcpyDeviceToHost,
                            C->stream));
      HIPCHK(hipStreamSynchronize(C->stream));
      const double cmax = *std::max_element(bm.begin(), bm.end());
      // largest diagonal near 2^14: far from the FP16 limit (65504), with the
      // smallest couplings (density ratio 1000) still normal numbers
      lv.scale16 = cmax > 0.0 ? std::ldexp(1.0, 14 - (int)std::ceil(std::log2(cmax))) : 1.0;
      _Float16 **dst[4] = {&lv.c16, &lv.w16, &lv.s16, &lv.b16};
      const double *src[4] = {s.c, s.w, s.s, s.b};
      for (int q = 0; q < 4; q++) {
        *dst[q] = C->harena.alloc<_Float16>(N);
        if (q == 0 && lv.d)  // flux form: scale * d
          hipLaunchKernelGGL(k_flux_d<_Float16>, dim3(grid_blocks(N)), dim3(kBlock), 0,
                             C->stream, lv.g, s.mat(), *dst[q], lv.scale16);
        else
          hipLaunchKernelGGL(k_to_f16, dim3(grid_blocks(N)), dim3(kBlock), 0, C->stream,
                             src[q], *dst[q], lv.scale16, N, q == 0 ? 1 : 0);
      }
      HIPCHK(hipGetLastError());
    }
    lv.use16 = on;
    lv.scale = on ? lv.scale16 : 1.0;
  }
}

// ---------------------------------------------------------------------------
// V-cycle as a preconditioner.  On entry the level's first sweep e1 is in
// lv.e[0] (written by the CG step on level 0, by the restriction below it).
// Returns the buffer that holds z.
// ---------------------------------------------------------------------------
template <typename T>
T *vcycle(psb_ctx *C, Hier<T> &H, int l, DevState *st, int nu1, int nu2, T omega, T cfac,
          int done = 0, bool rho0 = true);

// K-cycle levels chosen from the hierarchy (psb_stop.ac_kcycle = -1).
// Halving every direction: every level but the finest and the coarsest with
// 16^3 to 64^3 cells.  Measured on MI355X (mixed solves to 1e-8): at 256^3
// levels 2-4 take 16 iterations and 18.7 ms against the V-cycle's 36 and
// 27.7 ms, with level 1 added 12 and 21.6 ms; at 512^3 levels 3-5 66
// iterations and 352 ms against 113 and 548 ms, levels 2-5 89 and 515 ms.
// Larger levels cost more per visit than the iterations they save; smaller
// ones are a few launches of a few microseconds each, doubled at every K
// level above them.  Halving x and y before z (a z-stretched grid): the
// first four coarse levels, where that coarsening leaves the correction
// weakest; on the 384x384x768 cylinder case 60 iterations and 709 ms against
// 535 and 3201 ms, and 368 iterations with the levels of the sizes above.
constexpr int kKcycleMinCells = 1 << 12, kKcycleMaxCells = 1 << 18, kKcycleSemiLevels = 4;

// whether level l's coarse correction is the K-cycle's two Krylov steps
template <typename T>
bool kcycle_level(const psb_ctx *C, const Hier<T> &H, int l) {
  const int nl = (int)H.L.size();
  if (l < 1 || l >= nl - 1) return false;
  if (C->kmask == -1) {
    if (C->coarse_rule == 1) return l <= kKcycleSemiLevels;
    return H.L[l].g.N >= kKcycleMinCells && H.L[l].g.N <= kKcycleMaxCells;
  }
  return (C->kmask >> l) & 1;
}

// the K-cycle levels' work arrays, allocated before any graph capture
template <typename T>
void kcycle_buffers(psb_ctx *C, Hier<T> &H) {
  const int nl = (int)H.L.size();
  for (int l = 0; l < nl; l++) {
    Level<T> &lv = H.L[l];
    if (!kcycle_level(C, H, l)) continue;
    // each one that is missing: the batched solver allocates some of them on
    // the same levels
    if (!lv.kc) lv.kc = C->harena.alloc<T>(lv.g.N);
    if (!lv.kv) lv.kv = C->harena.alloc<T>(lv.g.N);
    if (!lv.kz) lv.kz = C->harena.alloc<T>(lv.g.N);
    if (!lv.krt) lv.krt = C->harena.alloc<T>(lv.g.N);
    if (!lv.kst) lv.kst = C->harena.alloc<double>(4);
  }
}

// the outer CG's beta for this solve: flexible when asked (ac_flexible 1),
// or by default (-1) when some level of the hierarchy takes the K-cycle's
// Krylov steps, i.e. when the preconditioner is not a fixed linear operator
template <typename T>
void set_flex(psb_ctx *C, const Hier<T> &H) {
  bool any = false;
  for (int l = 0; l < (int)H.L.size(); l++) any = any || kcycle_level(C, H, l);
  C->flex = C->flex_req == 1 || (C->flex_req == -1 && any);
}

// k_kc_r fused with the next cycle's second pre-sweep: rt to lv.krt, the
// sweep's output to lv.e[1] (the level's matrix as jacobi() picks it)
template <typename T>
void kc_r_sweep(psb_ctx *C, const Level<T> &lv, T om0, T omega, int nprev, DevState *st) {
  const Grid &g = lv.g;
  const int nb = nblocks<T>(g);
  auto launch = [&](auto mt, auto flux, const auto &A) {
    using M = typename decltype(mt)::type;
    constexpr bool FLUX = decltype(flux)::value;
    with_v<T>(g, [&](auto vt) {
      constexpr int V = decltype(vt)::value;
      hipLaunchKernelGGL((k_kc_r_sweep<T, V, M, FLUX>), dim3(nb), dim3(kBlock), 0, C->stream, g,
                         A, lv.c, lv.kv, lv.r, lv.krt, lv.e[1], om0, omega, lv.kst, nprev,
                         C->work.partial, st);
    });
  };
  using FT = std::true_type;
  using FF = std::false_type;
  if constexpr (std::is_same<T, float>::value) {
    if (lv.h16() && lv.d) return launch(std::common_type<_Float16>(), FT(), lv.mat16());
    if (lv.h16()) return launch(std::common_type<_Float16>(), FF(), lv.mat16());
    if (lv.d) return launch(std::common_type<T>(), FT(), lv.matflux());
  }
  launch(std::common_type<T>(), FF(), lv.mat());
}

// level l's correction for its right-hand side lv.r (first sweep already in
// lv.e[0]) by two flexible-CG steps preconditioned by the cycle on l (their
// reductions summed by the kernels that use them, no finalize launches):
// c = B r, v = s A c, alpha1 = c.r / c.v; rt = r - alpha1 v (in lv.r);
// d = B rt; gamma = d.v, beta = d.s A d, alpha2 = d.rt / (beta - gamma^2 /
// c.v); z = (alpha1 - gamma alpha2 / c.v) c + alpha2 d
template <typename T>
T *kstep(psb_ctx *C, Hier<T> &H, int l, DevState *st, int nu1, int nu2, T omega, T cfac) {
  Level<T> &lv = H.L[l];
  const Grid &g = lv.g;
  if (!lv.kc || !lv.kv || !lv.kz || !lv.krt || !lv.kst)  // kcycle_buffers() was not called
    throw HipError{hipErrorInvalidValue};
  const T s = T(lv.scale);
  const int nb = nblocks<T>(g);
  T *c = vcycle<T>(C, H, l, st, nu1, nu2, omega, cfac);
  with_v<T>(g, [&](auto vt) {
    constexpr int V = decltype(vt)::value;
    if (lv.d)
      hipLaunchKernelGGL((k_kc_av<T, V, true>), dim3(nb), dim3(kBlock), 0, C->stream, g,
                         lv.matflux(), s, c, lv.r, lv.kc, lv.kv, st, C->work);
    else
      hipLaunchKernelGGL((k_kc_av<T, V, false>), dim3(nb), dim3(kBlock), 0, C->stream, g,
                         lv.mat(), s, c, lv.r, lv.kc, lv.kv, st, C->work);
  });
  // with two or more pre-sweeps the residual update and the next cycle's
  // second sweep are one launch (the cycle then starts at its third sweep);
  // the updated right-hand side lives in lv.krt while the cycle and the
  // dot products run, lv.r keeping the level's own for the next visit
  const bool fuse = nu1 >= 2 && C->kfuse;
  T *r0 = lv.r;
  if (fuse) {
    kc_r_sweep<T>(C, lv, T(omega / lv.scale), omega, nb, st);
    lv.r = lv.krt;
  } else {
    with_v<T>(g, [&](auto vt) {  // sums k_kc_av's partials itself
      constexpr int V = decltype(vt)::value;
      hipLaunchKernelGGL((k_kc_r<T, V>), dim3(nb), dim3(kBlock), 0, C->stream, g, lv.c, lv.kv,
                         lv.r, lv.e[0], T(omega / lv.scale), lv.kst, nb, C->work.partial, st);
    });
  }
  T *d = vcycle<T>(C, H, l, st, nu1, nu2, omega, cfac, fuse ? 1 : 0);
  with_v<T>(g, [&](auto vt) {
    constexpr int V = decltype(vt)::value;
    if (lv.d)
      hipLaunchKernelGGL((k_kc_dw<T, V, true>), dim3(nb), dim3(kBlock), 0, C->stream, g,
                         lv.matflux(), s, d, lv.kv, lv.r, st, C->work);
    else
      hipLaunchKernelGGL((k_kc_dw<T, V, false>), dim3(nb), dim3(kBlock), 0, C->stream, g,
                         lv.mat(), s, d, lv.kv, lv.r, st, C->work);
  });
  with_v<T>(g, [&](auto vt) {  // sums k_kc_dw's partials itself
    constexpr int V = decltype(vt)::value;
    hipLaunchKernelGGL((k_kc_comb<T, V>), dim3(nb), dim3(kBlock), 0, C->stream, g, lv.kc, d,
                       lv.kz, lv.kst, nb, C->work.partial, st);
  });
  lv.r = r0;
  return lv.kz;
}

// done: pre-sweeps already applied beyond the first (whose result is in
// lv.e[done]); rho0: level 0's last post-sweep forms the CG's rho partials
// (false: a plain sweep, for a Krylov method with other scalars)
template <typename T>
T *vcycle(psb_ctx *C, Hier<T> &H, int l, DevState *st, int nu1, int nu2, T omega, T cfac,
          int done, bool rho0) {
  Level<T> &lv = H.L[l];
  const int nl = (int)H.L.size();
  if (l == nl - 1) {
    hipLaunchKernelGGL((k_gemv<T>), dim3((H.m + kWaves - 1) / kWaves), dim3(kBlock), 0,
                       C->stream, H.m, H.G, lv.r, lv.e[1], st);
    return lv.e[1];
  }
  int cur = done;
  for (int s = 1 + done; s < nu1; s++) {
    jacobi<T, false>(C, lv, lv.e[cur], lv.e[(cur + 1) % 3], nullptr, nullptr, T(0), omega, st);
    cur = (cur + 1) % 3;
  }
  Level<T> &cL = H.L[l + 1];
  // corrections on a level with a scaled half-precision matrix are held
  // divided by its scale: the coarse level's first sweep, the prolongation
  // and the preconditioner's output carry the ratios
  restrict_<T>(C, lv, lv.e[cur], cL, T(omega / cL.scale), st);
  const bool kc = kcycle_level(C, H, l + 1);  // the Krylov steps choose the scaling
  T *zH = kc ? kstep<T>(C, H, l + 1, st, nu1, nu2, omega, cfac)
             : vcycle<T>(C, H, l + 1, st, nu1, nu2, omega, cfac);
  const T cfac_l = T((kc ? 1.0 : (double)cfac) * (cL.scale / lv.scale));
  if (nu2 == 2) {  // both post-sweeps in one pass where the level suits it
    const int out = (cur + 2) % 3;
    const bool done = l == 0 && rho0
                          ? post2<T, true>(C, lv, lv.e[cur], lv.e[out], cL, zH, cfac_l, omega,
                                           T(lv.scale), st, H.qf)
                             : post2<T, false>(C, lv, lv.e[cur], lv.e[out], cL, zH, cfac_l,
                                               omega, T(1), st);
    if (done) return lv.e[out];
  }
  for (int s = 0; s < nu2; s++) {
    const int nxt = (cur + 1) % 3;
    const bool last = (s == nu2 - 1) && l == 0 && rho0;
    const Level<T> *co = s == 0 ? &cL : nullptr;
    const T *eH = s == 0 ? zH : nullptr;
    if (last)
      jacobi<T, true>(C, lv, lv.e[cur], lv.e[nxt], co, eH, cfac_l, omega, st, T(lv.scale),
                      H.qf);
    else
      jacobi<T, false>(C, lv, lv.e[cur], lv.e[nxt], co, eH, cfac_l, omega, st);
    cur = nxt;
  }
  return lv.e[cur];
}

// the CG p/q step on the level-0 FP16 copy (psb_stop.ac_store = 2): q =
// (scale A16) p / scale with FP32 arithmetic; false where it does not apply
template <typename T>
bool cg_pq16(psb_ctx *C, const Level<T> &f, const T *z, const T *pold, T *pnew, T *q,
             DevState *st) {
  if constexpr (std::is_same<T, float>::value) {
    MarchShape m;
    if (!C->cg_half || !f.h16() || !f.d || !vec_ok<float>(f.g) || !march_ok<4>(f.g, m))
      return false;
    {
      ProbeScope ps(C, PSB_K_CG_PQ, f.g, 4 * 2 + 4 * 4 + (f.aflag ? 1 : 0));
      hipLaunchKernelGGL((k_cg_pq_march<float, 4, true, false, true, _Float16>),
                         dim3(m.nblocks), dim3(kBlock), 0, C->stream, f.g, f.mat16(), z, pold,
                         pnew, q, f.aflag, st, C->work, m.TI, m.ntiles, (float)(1.0 / f.scale));
    }
    finalize<1>(C, FIN_PQ, m.nblocks, st);
    return true;
  } else {
    return false;
  }
}

// one preconditioned CG iteration: V-cycle (with rho), p/q step, x/r step
template <typename T>
void ac_iteration(psb_ctx *C, Hier<T> &H, KrylovVecs<T> &K, DevState *st, int parity, int nu1,
                  int nu2, T omega, T cfac, int xr_op = -1) {
  Level<T> &f = H.L[0];
  H.qf = C->flex ? K.q : nullptr;  // q = A p of the last direction (flexible beta)
  T *z = vcycle<T>(C, H, 0, st, nu1, nu2, omega, cfac);
  T *pold = K.p[parity], *pnew = K.p[parity ^ 1];
  if (!cg_pq16<T>(C, f, z, pold, pnew, K.q, st)) {
    if (f.d)
      cg_pq<T, true, false, true>(C, f.g, f.matflux(), z, pold, pnew, K.q, f.aflag, st);
    else
      cg_pq<T, true, false>(C, f.g, f.mat(), z, pold, pnew, K.q, f.aflag, st);
  }
  cg_xr<T, false>(C, f.g, f.c, pnew, K.q, K.x, K.r, f.e[0], T(omega / f.scale), st, xr_op);
}

void destroy_graph(Graph &gr) {
  if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
  if (gr.graph) (void)hipGraphDestroy(gr.graph);
  gr.exec = nullptr;
  gr.graph = nullptr;
  gr.key = -1;
  gr.fkey = 0.0;
  gr.px = gr.pb = nullptr;
}

// every captured graph holds device pointers of the matrix, hierarchy and
// geometry: any change to those drops them.  The device is idle first, so
// that no launch of a graph is in flight when it goes.
void destroy_graphs(psb_ctx *C) {
  if (C->stream) (void)hipStreamSynchronize(C->stream);
  (void)hipDeviceSynchronize();
  destroy_graph(C->gr_cgdiag);
  destroy_graph(C->gr_cgsgs);
  destroy_graph(C->gr_bicgs);
  for (auto &kv : C->gcache) destroy_graph(kv.second);
  C->gcache.clear();
  destroy_graph(C->href.gr);
  destroy_graph(C->href.gfull);
}

// the cached graph of an AC-preconditioned solve with these settings (empty
// if none was captured yet); the least recently used one goes when the
// cache is full
Graph &graph_for(psb_ctx *C, const GKey &k) {
  auto it = C->gcache.find(k);
  if (it == C->gcache.end()) {
    if (C->gcache.size() >= kGraphCacheMax) {
      auto old = C->gcache.begin();
      for (auto q = C->gcache.begin(); q != C->gcache.end(); ++q)
        if (q->second.used < old->second.used) old = q;
      HIPCHK(hipStreamSynchronize(C->stream));
      destroy_graph(old->second);
      C->gcache.erase(old);
    }
    it = C->gcache.emplace(k, Graph()).first;
  }
  it->second.used = ++C->gclock;
  return it->second;
}

// drop the hierarchy (and the graphs that use it): next AC solve rebuilds it
void reset_hier(psb_ctx *C) {
  HIPCHK(hipStreamSynchronize(C->stream));
  destroy_graphs(C);
  C->harena.release();
  C->h64 = Hier<double>();
  C->h32 = Hier<float>();
  C->href = RefHier();
}

// the hierarchy for the coarsening rule `rule` (the FP16 copies are added to
// it when a solve first asks for them, set_half)
void ensure_rule(psb_ctx *C, int rule) {
  if (rule == C->coarse_rule) return;
  reset_hier(C);
  C->coarse_rule = rule;
}


template <typename F>
void capture(psb_ctx *C, Graph &gr, long long key, F body) {
  if (gr.exec && gr.key == key) return;
  destroy_graph(gr);
  HIPCHK(hipStreamBeginCapture(C->stream, hipStreamCaptureModeThreadLocal));
  body();
  hipError_t e = hipStreamEndCapture(C->stream, &gr.graph);
  if (e != hipSuccess) throw HipError{e};
  HIPCHK(hipGraphInstantiate(&gr.exec, gr.graph, nullptr, nullptr, 0));
  gr.key = key;
}

void read_state(psb_ctx *C, DevState *st) {
  HIPCHK(hipMemcpyAsync(C->hst, st, sizeof(DevState), hipMemcpyDeviceToHost, C->stream));
  HIPCHK(hipStreamSynchronize(C->stream));
}

void init_state(psb_ctx *C, DevState *st, const psb_stop *p, int norm, double res_rat,
                double res_tol, int maxit, bool record = true) {
  DevState h;
  std::memset(&h, 0, sizeof(h));
  h.res_rat = res_rat;
  h.res_tol = res_tol;
  h.scale = p->scale;
  h.ntot = (double)C->g.N;
  h.mini = p->min_iter;
  h.maxit = maxit;
  h.stalecount = p->stalecount;  // at most kStaleMax (checked by the entry points)
  h.norm = norm;
  h.hist = record ? C->hist : nullptr;
  h.hist_cap = record ? C->hist_cap : 0;
  h.keepmean = C->singular ? 0 : 1;
  HIPCHK(hipMemcpyAsync(st, &h, sizeof(h), hipMemcpyHostToDevice, C->stream));
  HIPCHK(hipStreamSynchronize(C->stream));
}

void ensure_hist(psb_ctx *C, int maxit) {
  if (C->hist_cap >= maxit + 2) return;
  int cap = std::max(maxit + 2, 1024);
  C->hist = C->base.alloc<double>(cap);
  C->hist_cap = cap;
}

template <typename T>
void zero(psb_ctx *C, T *x, int n) {
  hipLaunchKernelGGL((k_fill<T>), dim3(grid_blocks(n)), dim3(kBlock), 0, C->stream, x, T(0), n);
}

void finish_result(psb_result *R, const DevState &h, int outer) {
  R->outer = outer;
  R->converged = h.converged;
  R->res0 = h.res0;
  R->res = h.res;
  R->iterations = h.report_iters;
  switch (h.reason) {
    case 1: case 2: case 7: R->status = PSB_OK; break;
    case 3: R->status = PSB_STALE; break;
    case 4: case 5: case 6: R->status = PSB_BREAKDOWN; break;
    case 8: R->status = PSB_NAN; break;
    default: R->status = h.converged ? PSB_OK : PSB_MAXITER;
  }
}

// true ||b - A x||_2 / ||b||_2 in FP64 (diagnostic, after the solve)
double true_rel(psb_ctx *C, const double *b, const double *x) {
  const Grid &g = C->g;
  auto go = [&](auto symtag) {
    constexpr bool S = decltype(symtag)::value;
    with_v<double>(g, [&](auto vt) {
      constexpr int V = decltype(vt)::value;
      hipLaunchKernelGGL((k_relres<V, S>), dim3(nblocks<double>(g)), dim3(kBlock), 0, C->stream,
                         g, mat64(C), b, x, C->work);
    });
  };
  if (C->sym) go(std::true_type());
  else go(std::false_type());
  finalize<2>(C, FIN_RELRES, nblocks<double>(g), C->st64, nullptr, C->relres);
  double h[2];
  HIPCHK(hipMemcpyAsync(h, C->relres, 16, hipMemcpyDeviceToHost, C->stream));
  HIPCHK(hipStreamSynchronize(C->stream));
  return h[1] > 0 ? std::sqrt(h[0] / h[1]) : 0.0;
}

template <typename F>
void run_graph(psb_ctx *C, Graph &gr, long long key, DevState *st, int maxit, F body) {
  long long launches = 0;
  while (!C->hst->done) {
    if (C->probe.on) {  // the same launches, one by one (timed per kernel)
      body();
    } else {
      capture(C, gr, key, body);
      HIPCHK(hipGraphLaunch(gr.exec, C->stream));
    }
    read_state(C, st);
    if (++launches > (long long)maxit + 2) break;
  }
}

int chunk_of(const psb_stop *p) { return std::max(2, p->graph_chunk + (p->graph_chunk & 1)); }

// ---------------------------------------------------------------------------
// reference CG with diagonal preconditioning (cg.cpp:10-182), FP64
// ---------------------------------------------------------------------------
template <bool SYM>
void solve_cg_diag(psb_ctx *C, double *x, const double *b, const psb_stop *p, psb_result *R) {
  KrylovVecs<double> &K = C->k64;
  ensure_krylov(C, K, false);
  ensure_hist(C, p->max_iter);
  const Grid &g = C->g;
  DMat<double> A = mat64(C);
  HIPCHK(hipMemcpyAsync(K.x, x, (size_t)g.N * 8, hipMemcpyDeviceToDevice, C->stream));
  init_state(C, C->st64, p, p->norm, p->res_rat, p->res_tol, p->max_iter);
  zero(C, K.p[0], g.N);
  zero(C, K.p[1], g.N);
  start<double, SYM, 0>(C, g, A, b, K.x, K.r, nullptr, 0.0, C->st64);
  HIPCHK(hipGetLastError());
  read_state(C, C->st64);
  const int chunk = chunk_of(p);
  auto body = [&]() {
    for (int it = 0; it < chunk; it++) {
      const int par = it & 1;
      cg_pq<double, SYM, true>(C, g, A, K.r, K.p[par], K.p[par ^ 1], K.q, nullptr, C->st64);
      cg_xr<double, true>(C, g, A.c, K.p[par ^ 1], K.q, K.x, K.r, nullptr, 0.0, C->st64);
    }
  };
  run_graph(C, C->gr_cgdiag, chunk * 2 + (SYM ? 1 : 0), C->st64, p->max_iter, body);
  HIPCHK(hipMemcpyAsync(x, K.x, (size_t)g.N * 8, hipMemcpyDeviceToDevice, C->stream));
  finish_result(R, *C->hst, 0);
  C->hist_count = std::min(C->hst->reason == 7 ? 1 : C->hst->iter + 1, C->hist_cap);
}

// ---------------------------------------------------------------------------
// CG with the diagonal preconditioner without any allocation
// (psb_solve_device_static): the kernels of solve_cg_diag launched eagerly
// (no graph: instantiating one allocates), iterating on the caller's x, with
// the Krylov vectors r, p0, p1, q, the reduction partials and the solver
// state in the file-scope device array `scratch` instead of buffers the
// library allocates.  Each device holds one such array for the life of the
// process, shared by every context and call on that device: calls that
// overlap on one device race.  Every region is written before it is read
// (the state by init_state, p0 / p1 cleared, r by the start kernel, q and
// the partials by their producing kernels), so a call never sees the
// previous call's contents.
// ---------------------------------------------------------------------------
constexpr size_t kStaticMaxCells = (size_t)1 << 21;  // 128^3 cells
constexpr size_t kStaticStateDoubles = (sizeof(DevState) + 255) / 256 * 32;
constexpr size_t kStaticPartialDoubles = (size_t)kMaxBlocks * kMaxPartials;
constexpr size_t kStaticDoubles =
    kStaticStateDoubles + kStaticPartialDoubles + 4 * kStaticMaxCells;  // 64.1 MiB

}  // namespace host
}  // namespace psb

static __device__ __align__(256) double scratch[psb::host::kStaticDoubles];

namespace psb {
namespace host {

// the device address of `scratch` on the current device, looked up once per
// device (the host-side name is not a device address)
double *static_scratch_here() {
  static std::mutex mu;
  static std::map<int, double *> ptrs;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto it = ptrs.find(dev);
  if (it != ptrs.end()) return it->second;
  void *ptr = nullptr;
  HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  ptrs[dev] = static_cast<double *>(ptr);
  return static_cast<double *>(ptr);
}

template <bool SYM>
void solve_cg_diag_static(psb_ctx *C, double *x, const double *b, const psb_stop *p,
                          psb_result *R) {
  const Grid &g = C->g;
  double *base = C->static_scratch;
  DevState *st = reinterpret_cast<DevState *>(base);
  double *partial = base + kStaticStateDoubles;
  double *r = partial + kStaticPartialDoubles;
  double *pv[2] = {r + kStaticMaxCells, r + 2 * kStaticMaxCells};
  double *q = r + 3 * kStaticMaxCells;
  // the launch helpers take the reduction partials from C->work
  struct WorkSwap {
    psb_ctx *C;
    Work keep;
    ~WorkSwap() { C->work = keep; }
  } swap{C, C->work};
  C->work.partial = partial;
  DMat<double> A = mat64(C);
  init_state(C, st, p, p->norm, p->res_rat, p->res_tol, p->max_iter, false);
  zero(C, pv[0], g.N);
  zero(C, pv[1], g.N);
  start<double, SYM, 0>(C, g, A, b, x, r, nullptr, 0.0, st);
  HIPCHK(hipGetLastError());
  read_state(C, st);
  const int chunk = chunk_of(p);
  long long rounds = 0;
  while (!C->hst->done) {
    for (int it = 0; it < chunk; it++) {
      const int par = it & 1;
      cg_pq<double, SYM, true>(C, g, A, r, pv[par], pv[par ^ 1], q, nullptr, st);
      cg_xr<double, true>(C, g, A.c, pv[par ^ 1], q, x, r, nullptr, 0.0, st);
    }
    HIPCHK(hipGetLastError());
    read_state(C, st);
    if (++rounds > (long long)p->max_iter + 2) break;
  }
  finish_result(R, *C->hst, 0);
  C->hist_count = 0;
  R->rel_l2 = true_rel(C, b, x);
}

// ---------------------------------------------------------------------------
// CG (cg.cpp:10-182) with the red-black symmetric Gauss-Seidel preconditioner
// in place of Prec::ic0 / ic2 / ic3, whose triangular sweeps
// (preconditioner_solve.cpp:14-110) run cell after cell: per iteration
// z = r / c (fused into the x/r step), the two colour sweeps (the second
// with rho = r.z), the p/q step and the x/r step with the stop test.  FP64.
// ---------------------------------------------------------------------------
template <bool SYM>
void rbsgs(psb_ctx *C, const Grid &g, DMat<double> A, double *z, const double *r, int colour,
           bool rho, DevState *st) {
  with_v<double>(g, [&](auto vt) {
    constexpr int V = decltype(vt)::value;
    if (rho)
      hipLaunchKernelGGL((k_rbsgs<double, V, SYM, true>), dim3(nblocks<double>(g)), dim3(kBlock),
                         0, C->stream, g, A, z, r, colour, st, C->work);
    else
      hipLaunchKernelGGL((k_rbsgs<double, V, SYM, false>), dim3(nblocks<double>(g)),
                         dim3(kBlock), 0, C->stream, g, A, z, r, colour, st, C->work);
  });
  if (rho) finalize<4>(C, FIN_RHO, nblocks<double>(g), st);
}

template <bool SYM>
void solve_cg_sgs(psb_ctx *C, double *x, const double *b, const psb_stop *p, psb_result *R) {
  KrylovVecs<double> &K = C->k64;
  ensure_krylov(C, K, false);
  if (!K.z) K.z = C->arena.alloc<double>(C->g.N);
  ensure_hist(C, p->max_iter);
  const Grid &g = C->g;
  DMat<double> A = mat64(C);
  HIPCHK(hipMemcpyAsync(K.x, x, (size_t)g.N * 8, hipMemcpyDeviceToDevice, C->stream));
  init_state(C, C->st64, p, p->norm, p->res_rat, p->res_tol, p->max_iter);
  zero(C, K.p[0], g.N);
  zero(C, K.p[1], g.N);
  start<double, SYM, 1>(C, g, A, b, K.x, K.r, K.z, 1.0, C->st64);  // z = r / c
  HIPCHK(hipGetLastError());
  read_state(C, C->st64);
  const int chunk = chunk_of(p);
  auto body = [&]() {
    for (int it = 0; it < chunk; it++) {
      const int par = it & 1;
      rbsgs<SYM>(C, g, A, K.z, K.r, 1, false, C->st64);
      rbsgs<SYM>(C, g, A, K.z, K.r, 0, true, C->st64);
      cg_pq<double, SYM, false>(C, g, A, K.z, K.p[par], K.p[par ^ 1], K.q, nullptr, C->st64);
      cg_xr<double, false>(C, g, A.c, K.p[par ^ 1], K.q, K.x, K.r, K.z, 1.0, C->st64);
    }
  };
  run_graph(C, C->gr_cgsgs, chunk * 2 + (SYM ? 1 : 0), C->st64, p->max_iter, body);
  HIPCHK(hipMemcpyAsync(x, K.x, (size_t)g.N * 8, hipMemcpyDeviceToDevice, C->stream));
  finish_result(R, *C->hst, 0);
  C->hist_count = std::min(C->hst->reason == 7 ? 1 : C->hst->iter + 1, C->hist_cap);
}

// ---------------------------------------------------------------------------
// reference BiCGS with diagonal preconditioning (bicgs.cpp:9-229), FP64
// ---------------------------------------------------------------------------
template <bool SYM>
void solve_bicgs_diag(psb_ctx *C, double *x, const double *b, const psb_stop *p,
                      psb_result *R) {
  KrylovVecs<double> &K = C->k64;
  ensure_krylov(C, K, true);
  ensure_hist(C, p->max_iter);
  const Grid &g = C->g;
  DMat<double> A = mat64(C);
  init_state(C, C->st64, p, p->norm, p->res_rat, p->res_tol, p->max_iter);
  HIPCHK(hipMemcpyAsync(K.x, x, (size_t)g.N * 8, hipMemcpyDeviceToDevice, C->stream));
  for (int q = 0; q < 2; q++) {
    zero(C, K.p[q], g.N);
    zero(C, K.v[q], g.N);
  }
  start<double, SYM, 2>(C, g, A, b, K.x, K.r, K.rt, 0.0, C->st64);
  HIPCHK(hipGetLastError());
  read_state(C, C->st64);
  const int chunk = chunk_of(p);
  const int nb = nblocks<double>(g);
  auto body = [&]() {
    for (int it = 0; it < chunk; it++) {
      const int par = it & 1;
      with_v<double>(g, [&](auto vt) {
        constexpr int V = decltype(vt)::value;
        hipLaunchKernelGGL((k_bi1<double, V, SYM>), dim3(nb), dim3(kBlock), 0, C->stream, g, A,
                           K.r, K.rt, K.p[par], K.v[par], K.p[par ^ 1], K.v[par ^ 1], C->st64,
                           C->work);
        finalize<1>(C, FIN_BI1, nb, C->st64);
        hipLaunchKernelGGL((k_bi2<double, V, SYM>), dim3(nb), dim3(kBlock), 0, C->stream, g, A,
                           K.r, K.v[par ^ 1], K.s, K.t, C->st64, C->work);
        finalize<3>(C, FIN_BI2, nb, C->st64);
        hipLaunchKernelGGL((k_bi3<double, V>), dim3(nb), dim3(kBlock), 0, C->stream, g, A.c,
                           metric(C), K.p[par ^ 1], K.s, K.t, K.rt, K.x, K.r, C->st64, C->work);
        finalize<3>(C, FIN_BI3, nb, C->st64);
      });
    }
  };
  run_graph(C, C->gr_bicgs, chunk * 2 + (SYM ? 1 : 0), C->st64, p->max_iter, body);
  HIPCHK(hipMemcpyAsync(x, K.x, (size_t)g.N * 8, hipMemcpyDeviceToDevice, C->stream));
  finish_result(R, *C->hst, 0);
  C->hist_count = std::min(C->hst->reason == 7 ? 1 : C->hst->iter + 1, C->hist_cap);
}

// ---------------------------------------------------------------------------
// AC-preconditioned CG in precision T on K.x, K.r (r = b - A x already
// formed, e1 in level 0's e[0], st initialised).  Runs to st->done.
// ---------------------------------------------------------------------------
// the graph key of an AC-preconditioned solve with the context's current
// settings (body: 0 FP64, 1 FP32 inner solve, 2 reliable updates, 3 FP64
// BiCGS, 4 BiCGS with the FP32 cycle)
GKey solve_key(const psb_ctx *C, const psb_stop *p, int body, const void *x, const void *b) {
  GKey k;
  k.body = body;
  k.chunk = chunk_of(p);
  k.nu1 = p->ac_pre;
  k.nu2 = p->ac_post;
  k.kmask = C->kmask;
  k.fuse = C->fuse;
  k.cg16 = C->cg_half;
  k.half = body == 0 || body >= 3 ? 0 : (C->h32.L.empty() ? 0 : C->h32.L[0].use16);
  k.flex = C->flex;
  k.omega = p->ac_omega;
  k.cfac = p->ac_cfac;
  k.px = x;
  k.pb = b;
  return k;
}

template <typename T>
void run_acpcg(psb_ctx *C, Hier<T> &H, KrylovVecs<T> &K, DevState *st, int body,
               const psb_stop *p) {
  const int chunk = chunk_of(p);
  const int nu1 = p->ac_pre, nu2 = p->ac_post;
  const T omega = (T)p->ac_omega, cfac = (T)p->ac_cfac;
  auto f = [&]() {
    for (int it = 0; it < chunk; it++) ac_iteration<T>(C, H, K, st, it & 1, nu1, nu2, omega, cfac);
  };
  Graph &gr = graph_for(C, solve_key(C, p, body, nullptr, nullptr));
  run_graph(C, gr, 0, st, p->max_iter, f);
}

void solve_ac64(psb_ctx *C, double *x, const double *b, const psb_stop *p, psb_result *R) {
  KrylovVecs<double> &K = C->k64;
  ensure_krylov(C, K, false);
  build_hier64(C);
  kcycle_buffers(C, C->h64);
  set_flex(C, C->h64);
  ensure_hist(C, p->max_iter);
  const Grid &g = C->g;
  Level<double> &f = C->h64.L[0];
  // the graph works on K.x; copy the caller's x in and out
  HIPCHK(hipMemcpyAsync(K.x, x, (size_t)g.N * 8, hipMemcpyDeviceToDevice, C->stream));
  init_state(C, C->st64, p, p->norm, p->res_rat, p->res_tol, p->max_iter);
  zero(C, K.p[0], g.N);
  zero(C, K.p[1], g.N);
  start<double, true, 1>(C, g, mat64(C), b, K.x, K.r, f.e[0], p->ac_omega, C->st64);
  HIPCHK(hipGetLastError());
  read_state(C, C->st64);
  run_acpcg<double>(C, C->h64, K, C->st64, 0, p);
  HIPCHK(hipMemcpyAsync(x, K.x, (size_t)g.N * 8, hipMemcpyDeviceToDevice, C->stream));
  finish_result(R, *C->hst, 0);
  C->hist_count = std::min(C->hst->iter + 1, C->hist_cap);
}

// ---------------------------------------------------------------------------
// BiCGS preconditioned by the multigrid V-cycle, FP64 (psb_stop.method =
// PSB_BICGS with PSB_PREC_AC): the steps of bicgs.cpp:79-229 with the cycle
// as M^-1, applied to p and to s (level 0's right-hand side pointed at each
// in turn, its first sweep formed by the pass before).  A non-symmetric
// matrix keeps 7 arrays on every level (span sweeps and restriction, the
// coarsest inverse by LU).  The K-cycle's steps
// change the preconditioner from one application to the next, which
// BiCGStab does not allow for, so the levels run the V-cycle whatever
// ac_kcycle says.
// ---------------------------------------------------------------------------
template <bool SYM>
void solve_acbicgs64(psb_ctx *C, double *x, const double *b, const psb_stop *p,
                     psb_result *R) {
  KrylovVecs<double> &K = C->k64;
  ensure_krylov(C, K, true);
  if (!K.z) K.z = C->arena.alloc<double>(C->g.N);
  build_hier64(C);
  ensure_hist(C, p->max_iter);
  Hier<double> &H = C->h64;
  const Grid &g = C->g;
  const DMat<double> A = mat64(C);
  HIPCHK(hipMemcpyAsync(K.x, x, (size_t)g.N * 8, hipMemcpyDeviceToDevice, C->stream));
  init_state(C, C->st64, p, p->norm, p->res_rat, p->res_tol, p->max_iter);
  for (int q = 0; q < 2; q++) {
    zero(C, K.p[q], g.N);
    zero(C, K.v[q], g.N);
  }
  start<double, SYM, 2>(C, g, A, b, K.x, K.r, K.rt, 0.0, C->st64);
  HIPCHK(hipGetLastError());
  read_state(C, C->st64);
  const int chunk = chunk_of(p);
  const int nb = nblocks<double>(g);
  const int nu1 = p->ac_pre, nu2 = p->ac_post;
  const double omega = p->ac_omega, cfac = p->ac_cfac;
  DevState *st = C->st64;
  H.qf = nullptr;
  auto body = [&]() {
    Level<double> &f = H.L[0];
    double *r0 = f.r;
    const double omj = omega / f.scale;
    for (int it = 0; it < chunk; it++) {
      const int par = it & 1;
      with_v<double>(g, [&](auto vt) {
        constexpr int V = decltype(vt)::value;
        hipLaunchKernelGGL((k_abi_p<double, V>), dim3(nb), dim3(kBlock), 0, C->stream, g, f.c,
                           K.r, K.p[par], K.v[par], K.p[par ^ 1], f.e[0], omj, st);
      });
      f.r = K.p[par ^ 1];
      const double *z1 = vcycle<double>(C, H, 0, st, nu1, nu2, omega, cfac, 0, false);
      with_v<double>(g, [&](auto vt) {
        constexpr int V = decltype(vt)::value;
        hipLaunchKernelGGL((k_abi_v<double, V, SYM>), dim3(nb), dim3(kBlock), 0, C->stream, g,
                           A, z1, K.rt, K.v[par ^ 1], K.z, st, C->work);
        finalize<1>(C, FIN_BI1, nb, st);
        hipLaunchKernelGGL((k_abi_s<double, V>), dim3(nb), dim3(kBlock), 0, C->stream, g, f.c,
                           K.r, K.v[par ^ 1], K.s, f.e[0], omj, st);
      });
      f.r = K.s;
      const double *z2 = vcycle<double>(C, H, 0, st, nu1, nu2, omega, cfac, 0, false);
      with_v<double>(g, [&](auto vt) {
        constexpr int V = decltype(vt)::value;
        hipLaunchKernelGGL((k_abi_t<double, V, SYM>), dim3(nb), dim3(kBlock), 0, C->stream, g,
                           A, z2, K.s, K.t, st, C->work);
        finalize<3>(C, FIN_BI2, nb, st);
        hipLaunchKernelGGL((k_abi_x<double, V>), dim3(nb), dim3(kBlock), 0, C->stream, g,
                           metric(C), K.z, z2, K.s, K.t, K.rt, K.x, K.r, st, C->work);
        finalize<3>(C, FIN_BI3, nb, st);
      });
      f.r = r0;
    }
  };
  Graph &gr = graph_for(C, solve_key(C, p, 3, nullptr, nullptr));
  run_graph(C, gr, 0, st, p->max_iter, body);
  HIPCHK(hipMemcpyAsync(x, K.x, (size_t)g.N * 8, hipMemcpyDeviceToDevice, C->stream));
  finish_result(R, *C->hst, 0);
  C->hist_count = std::min(C->hst->reason == 7 ? 1 : C->hst->iter + 1, C->hist_cap);
}

// ---------------------------------------------------------------------------
// the same BiCGS with the V-cycle on the FP32 levels (psb_stop.precision =
// PSB_MIXED, symmetric matrices): p and s go to FP32 with the level-0 first
// sweep, the cycle's output comes back to FP64; the BiCGS vectors, products
// and scalars stay FP64.  The FP32 levels' smoother reads FP32 matrices
// (no half-precision copies: the cycle's output is then unscaled).
// ---------------------------------------------------------------------------
void solve_acbicgs_mixed(psb_ctx *C, double *x, const double *b, const psb_stop *p,
                         psb_result *R) {
  KrylovVecs<double> &K = C->k64;
  ensure_krylov(C, K, true);
  if (!K.z) K.z = C->arena.alloc<double>(C->g.N);
  build_hier32(C);
  set_half(C, false);
  ensure_krylov(C, C->k32, true);
  ensure_hist(C, p->max_iter);
  Hier<float> &H = C->h32;
  const Grid &g = C->g;
  const DMat<double> A = mat64(C);
  HIPCHK(hipMemcpyAsync(K.x, x, (size_t)g.N * 8, hipMemcpyDeviceToDevice, C->stream));
  init_state(C, C->st64, p, p->norm, p->res_rat, p->res_tol, p->max_iter);
  for (int q = 0; q < 2; q++) {
    zero(C, K.p[q], g.N);
    zero(C, K.v[q], g.N);
  }
  start<double, true, 2>(C, g, A, b, K.x, K.r, K.rt, 0.0, C->st64);
  HIPCHK(hipGetLastError());
  read_state(C, C->st64);
  const int chunk = chunk_of(p);
  const int nb = nblocks<double>(g);
  const int nu1 = p->ac_pre, nu2 = p->ac_post;
  const float omega = (float)p->ac_omega, cfac = (float)p->ac_cfac;
  DevState *st = C->st64;
  H.qf = nullptr;
  double *z2 = K.q;  // the second cycle's output in FP64
  auto body = [&]() {
    Level<float> &f = H.L[0];
    float *r0 = f.r;
    for (int it = 0; it < chunk; it++) {
      const int par = it & 1;
      with_v<double>(g, [&](auto vt) {
        constexpr int V = decltype(vt)::value;
        hipLaunchKernelGGL((k_abi_p<double, V>), dim3(nb), dim3(kBlock), 0, C->stream, g, A.c,
                           K.r, K.p[par], K.v[par], K.p[par ^ 1], nullptr, 0.0, st);
        hipLaunchKernelGGL((k_abi_to32<V>), dim3(nb), dim3(kBlock), 0, C->stream, g,
                           K.p[par ^ 1], f.c, C->k32.p[0], f.e[0], omega, st);
      });
      f.r = C->k32.p[0];
      const float *z1 = vcycle<float>(C, H, 0, st, nu1, nu2, omega, cfac, 0, false);
      with_v<double>(g, [&](auto vt) {
        constexpr int V = decltype(vt)::value;
        hipLaunchKernelGGL((k_abi_to64<V>), dim3(nb), dim3(kBlock), 0, C->stream, g, z1, K.z,
                           st);
        hipLaunchKernelGGL((k_abi_v<double, V, true>), dim3(nb), dim3(kBlock), 0, C->stream, g,
                           A, K.z, K.rt, K.v[par ^ 1], nullptr, st, C->work);
        finalize<1>(C, FIN_BI1, nb, st);
        hipLaunchKernelGGL((k_abi_s<double, V>), dim3(nb), dim3(kBlock), 0, C->stream, g, A.c,
                           K.r, K.v[par ^ 1], K.s, nullptr, 0.0, st);
        hipLaunchKernelGGL((k_abi_to32<V>), dim3(nb), dim3(kBlock), 0, C->stream, g, K.s, f.c,
                           C->k32.s, f.e[0], omega, st);
      });
      f.r = C->k32.s;
      const float *z2f = vcycle<float>(C, H, 0, st, nu1, nu2, omega, cfac, 0, false);
      with_v<double>(g, [&](auto vt) {
        constexpr int V = decltype(vt)::value;
        hipLaunchKernelGGL((k_abi_to64<V>), dim3(nb), dim3(kBlock), 0, C->stream, g, z2f, z2,
                           st);
        hipLaunchKernelGGL((k_abi_t<double, V, true>), dim3(nb), dim3(kBlock), 0, C->stream, g,
                           A, z2, K.s, K.t, st, C->work);
        finalize<3>(C, FIN_BI2, nb, st);
        hipLaunchKernelGGL((k_abi_x<double, V>), dim3(nb), dim3(kBlock), 0, C->stream, g,
                           metric(C), K.z, z2, K.s, K.t, K.rt, K.x, K.r, st, C->work);
        finalize<3>(C, FIN_BI3, nb, st);
      });
      f.r = r0;
    }
  };
  Graph &gr = graph_for(C, solve_key(C, p, 4, nullptr, nullptr));
  run_graph(C, gr, 0, st, p->max_iter, body);
  HIPCHK(hipMemcpyAsync(x, K.x, (size_t)g.N * 8, hipMemcpyDeviceToDevice, C->stream));
  finish_result(R, *C->hst, 0);
  C->hist_count = std::min(C->hst->reason == 7 ? 1 : C->hst->iter + 1, C->hist_cap);
}

// ---------------------------------------------------------------------------
// mixed precision: FP64 iterative refinement around the FP32 AC-PCG
// (r = b - A x in FP64, inner solve on the centred r, x += d in FP64)
// ---------------------------------------------------------------------------
void solve_ir(psb_ctx *C, double *x, const double *b, const psb_stop *p, psb_result *R) {
  KrylovVecs<float> &K = C->k32;
  ensure_krylov(C, K, false);
  build_hier32(C);
  kcycle_buffers(C, C->h32);
  set_flex(C, C->h32);
  ensure_hist(C, p->max_iter);
  const Grid &g = C->g;
  Level<float> &f = C->h32.L[0];
  init_state(C, C->st64, p, p->norm, p->res_rat, p->res_tol, p->max_iter);
  psb_stop in = *p;
  in.norm = PSB_NORM_L2REL;
  in.min_iter = 1;
  in.stalecount = -1;
  init_state(C, C->st32, &in, PSB_NORM_L2REL, p->inner_rtol, 0.0, p->max_iter, false);
  zero(C, K.x, g.N);
  zero(C, K.p[0], g.N);
  zero(C, K.p[1], g.N);
  double res0 = -1.0, prev = 1e300;
  int outer = 0, inner_total = 0, status = PSB_MAXITER, conv = 0;
  C->hist_host.clear();
  const int nb64 = nblocks<double>(g), nb32 = nblocks<float>(g);
  for (;;) {
    with_v<double>(g, [&](auto vt) {
      constexpr int V = decltype(vt)::value;
      hipLaunchKernelGGL((k_ir_resid<V, true>), dim3(nb64), dim3(kBlock), 0, C->stream, g,
                         mat64(C), metric(C), b, x, K.r, f.aflag, C->work);
    });
    finalize<4>(C, FIN_IR_RESID, nb64, C->st64, C->st32);
    HIPCHK(hipGetLastError());
    DevState h64;
    HIPCHK(hipMemcpyAsync(&h64, C->st64, sizeof(h64), hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    const double res = h64.res;
    C->hist_host.push_back(res);
    if (res0 < 0) res0 = res;
    R->res = res;
    if (!(res == res)) { status = PSB_NAN; break; }
    if (res < p->res_tol || res < res0 * p->res_rat || res0 == 0.0) {
      conv = 1;
      status = PSB_OK;
      break;
    }
    if (outer >= p->max_outer) { status = PSB_MAXITER; break; }
    if (outer > 0 && res > 0.9 * prev) { status = PSB_DIVERGED; break; }
    prev = res;
    // relative reduction asked of this inner solve
    double rt = p->inner_rtol;
    if (p->inner_adapt) {
      const double target = std::max(p->res_tol, res0 * p->res_rat);
      rt = std::max(p->inner_rtol, std::min(0.5, 0.5 * target / res));
    }
    HIPCHK(hipMemcpyAsync(&C->st32->res_rat, &rt, sizeof(double), hipMemcpyHostToDevice,
                          C->stream));
    // inner FP32 solve on the centred residual
    with_v<float>(g, [&](auto vt) {
      constexpr int V = decltype(vt)::value;
      hipLaunchKernelGGL((k_center<V>), dim3(nb32), dim3(kBlock), 0, C->stream, g, f.c, f.aflag,
                         K.r, f.e[0], (float)(p->ac_omega / f.scale), C->st32, C->work);
    });
    finalize<1>(C, FIN_CENTER, nb32, C->st32);
    HIPCHK(hipGetLastError());
    read_state(C, C->st32);
    run_acpcg<float>(C, C->h32, K, C->st32, 1, &in);
    inner_total += C->hst->iter;
    if (C->hst->reason == 8) { status = PSB_NAN; break; }
    with_v<double>(g, [&](auto vt) {
      constexpr int V = decltype(vt)::value;
      hipLaunchKernelGGL((k_ir_update<V>), dim3(nb64), dim3(kBlock), 0, C->stream, g, x, K.x);
    });
    HIPCHK(hipGetLastError());
    outer++;
  }
  C->hist_on_host = true;
  C->hist_count = (int)C->hist_host.size();
  R->res0 = res0;
  R->outer = outer;
  R->iterations = inner_total;
  R->converged = conv;
  R->status = status;
}

// ---------------------------------------------------------------------------
// mixed precision with reliable updates: one FP32 AC-PCG run that is never
// restarted.  Its iterate d accumulates the correction since the last
// refresh; whenever its recursive residual has fallen by inner_rtol below the
// last true residual (or reaches the target), x += d in FP64 and the true
// residual b - A x, rounded to FP32, replaces the recursive one.  The search
// direction and rho carry over, so CG keeps its Krylov space instead of
// rebuilding it after every refinement step.
// (H. A. van der Vorst and Q. Ye, "Residual replacement strategies for
// Krylov subspace iterative methods", SIAM J. Sci. Comput. 22 (2000).)
// ---------------------------------------------------------------------------
// in a timed eager solve: whether a gated reliable-update kernel will run
// (the refresh flag its finalize left), read back from the device; outside
// one (or ungated) always true
bool probe_gate(psb_ctx *C, bool gated) {
  if (!C->probe.on || !gated) return true;
  read_state(C, C->st32);
  return C->hst->update != 0 && !C->hst->done;
}

void ru_resid(psb_ctx *C, const double *b, const double *x, bool gated, int op) {
  const Grid &g = C->g;
  Level<float> &f = C->h32.L[0];
  KrylovVecs<float> &K = C->k32;
  const float omega = (float)(C->ac_omega / f.scale);
  int nb = nblocks<double>(g);
  {
    // FP64 c, w, s, b, x, b and FP32 c in; FP32 r and e1 out (timed only
    // when it runs: a gated launch without a pending update returns at once)
    ProbeScope ps(C, PSB_K_RU_RESID, probe_gate(C, gated) ? g : Grid{}, 6 * 8 + 3 * 4);
    with_v<double>(g, [&](auto vt) {
      constexpr int V = decltype(vt)::value;
      MarchShape m;
      if (lds_use<double>(C, kLdsResid) && V > 1 && g.N >= kMarchMin && lds_ok<V>(g) &&
          march_ok<V>(g, m)) {
        nb = m.nblocks;
        hipLaunchKernelGGL((k_ru_resid_lds<V, true>), dim3(m.nblocks), dim3(kBlock), 0,
                           C->stream, g, mat64(C), metric(C), b, x, f.c, K.r, f.e[0], omega,
                           C->st32, gated, C->work, m.TI, m.ntiles);
        return;
      }
      hipLaunchKernelGGL((k_ru_resid<V, true>), dim3(nblocks<double>(g)), dim3(kBlock), 0,
                         C->stream, g, mat64(C), metric(C), b, x, f.c, K.r, f.e[0], omega,
                         C->st32, gated, C->work);
    });
  }
  finalize<4>(C, op, nb, C->st32);
}

void ru_xupdate(psb_ctx *C, double *x, bool gated) {
  const Grid &g = C->g;
  // x, d in and out (timed only when a refresh is pending)
  ProbeScope ps(C, PSB_K_RU_XUPDATE, probe_gate(C, gated) ? g : Grid{}, 2 * 8 + 2 * 4);
  with_v<double>(g, [&](auto vt) {
    constexpr int V = decltype(vt)::value;
    hipLaunchKernelGGL((k_ru_xupdate<V>), dim3(nblocks<double>(g)), dim3(kBlock), 0, C->stream,
                       g, x, C->k32.x, C->st32, gated);
  });
}

void solve_ru(psb_ctx *C, double *x, const double *b, const psb_stop *p, psb_result *R) {
  KrylovVecs<float> &K = C->k32;
  ensure_krylov(C, K, false);
  build_hier32(C);
  kcycle_buffers(C, C->h32);
  set_flex(C, C->h32);
  ensure_hist(C, p->max_iter);
  const Grid &g = C->g;
  const int chunk = chunk_of(p);
  const int nu1 = p->ac_pre, nu2 = p->ac_post;
  const float omega = (float)p->ac_omega, cfac = (float)p->ac_cfac;
  Graph &gr = graph_for(C, solve_key(C, p, 2, x, b));
  init_state(C, C->st32, p, p->norm, p->res_rat, p->res_tol, p->max_iter);
  HIPCHK(hipMemcpyAsync(&C->st32->delta, &p->inner_rtol, sizeof(double), hipMemcpyHostToDevice,
                        C->stream));
  zero(C, K.x, g.N);
  zero(C, K.p[0], g.N);
  zero(C, K.p[1], g.N);
  ru_resid(C, b, x, false, FIN_RU_START);
  HIPCHK(hipGetLastError());
  read_state(C, C->st32);
  auto body = [&]() {
    for (int it = 0; it < chunk; it++) {
      ac_iteration<float>(C, C->h32, K, C->st32, it & 1, nu1, nu2, omega, cfac, FIN_XR_RU);
      ru_xupdate(C, x, true);
      ru_resid(C, b, x, true, FIN_RU);
    }
  };
  run_graph(C, gr, 0, C->st32, p->max_iter, body);
  // fold in what the last FP32 iterations added since the last refresh
  ru_xupdate(C, x, false);
  DevState h = *C->hst;
  if (!h.converged && h.reason != 8) {  // report the final true residual
    ru_resid(C, b, x, false, FIN_RU_END);
    read_state(C, C->st32);
    h.res_true = C->hst->res_true;
    h.rl2_true = C->hst->rl2_true;
  }
  finish_result(R, h, h.outer);
  R->res = h.res_true;
  // ||b - A x|| / ||b|| from the last FP64 residual (the x it was taken at:
  // a converged solve stops right after a refresh, which leaves no FP32
  // correction behind), instead of another pass over the FP64 system
  R->rel_l2 = h.bsq > 0.0 ? h.rl2_true / std::sqrt(h.bsq) : 0.0;
  C->hist_count = std::min(h.outer + 1, C->hist_cap);
}

// ---------------------------------------------------------------------------
// the reference's AC::cycle with the V loop (additive_cycle.cpp:4-99,
// additive_recursive.cpp:4-34): on every level a CG with the diagonal
// preconditioner smooths (MaxIter mv[0] before and mv[2] after the coarse
// correction, ResRat and ResTol atto as AC's defaults, additive.cpp:16-17),
// the coarsest level gets 100 CG iterations, the residual is restricted with
// the mean over active cells removed, the correction interpolated; the coarse
// levels' phi persists between cycles of one call.
// ---------------------------------------------------------------------------
constexpr double kAtto = 1.0e-18;  // Src/Global/global_constants.cpp:19

void build_ref(psb_ctx *C) {
  RefHier &H = C->href;
  if (H.built) return;
  hipStream_t stream = C->stream;
  std::vector<int> d = level_dims(C->g.nx, C->g.ny, C->g.nz, 0, 2);
  const int nl = (int)d.size() / 3;
  if (nl < 2) throw std::invalid_argument("grid too small for the AC cycle");
  H.L.assign(nl, Level<double>());
  H.M.assign(nl, Metric{nullptr, nullptr, nullptr});
  H.K.assign(nl, std::vector<double *>(4, nullptr));
  std::vector<double> hx = C->gdx, hy = C->gdy, hz = C->gdz;
  if (hx.empty()) {
    hx.assign(C->g.nx, 1.0);
    hy.assign(C->g.ny, 1.0);
    hz.assign(C->g.nz, 1.0);
  }
  for (int l = 0; l < nl; l++) {
    Level<double> &lv = H.L[l];
    lv.g = make_grid(d[3 * l], d[3 * l + 1], d[3 * l + 2]);
    const size_t N = lv.g.N;
    if (l == 0) {
      lv.c = C->A64[0]; lv.w = C->A64[1]; lv.s = C->A64[2]; lv.b = C->A64[3];
      if (!C->sym) { lv.ae = C->A64[4]; lv.an = C->A64[5]; lv.at = C->A64[6]; }
      lv.aflag = C->aflag0;
    } else {
      Level<double> &f = H.L[l - 1];
      lv.CI = f.g.nx / lv.g.nx;
      lv.CJ = f.g.ny / lv.g.ny;
      lv.CK = f.g.nz / lv.g.nz;
      lv.c = C->harena.alloc<double>(N); lv.w = C->harena.alloc<double>(N);
      lv.s = C->harena.alloc<double>(N); lv.b = C->harena.alloc<double>(N);
      if (!C->sym) {  // a non-symmetric matrix keeps 7 arrays on every level
        lv.ae = C->harena.alloc<double>(N); lv.an = C->harena.alloc<double>(N);
        lv.at = C->harena.alloc<double>(N);
      }
      hipLaunchKernelGGL(k_galerkin, dim3(grid_blocks(N)), dim3(kBlock), 0, stream, f.g,
                         f.mat(), lv.g, lv.CI, lv.CJ, lv.CK, lv.c, lv.w, lv.s, lv.b, lv.ae,
                         lv.an, lv.at);
      if (f.aflag) {
        lv.aflag = C->harena.alloc<uint8_t>(N);
        hipLaunchKernelGGL(k_coarsen_flag, dim3(grid_blocks(N)), dim3(kBlock), 0, stream, f.g,
                           f.aflag, lv.g, lv.CI, lv.CJ, lv.CK, lv.aflag);
      }
      // coarse cell sizes: sums over the block (the reference's coarse grid
      // spans the same nodes)
      std::vector<double> cx(lv.g.nx, 0.0), cy(lv.g.ny, 0.0), cz(lv.g.nz, 0.0);
      for (int i = 0; i < f.g.nx; i++) cx[i / lv.CI] += hx[i];
      for (int j = 0; j < f.g.ny; j++) cy[j / lv.CJ] += hy[j];
      for (int k = 0; k < f.g.nz; k++) cz[k / lv.CK] += hz[k];
      hx = cx;
      hy = cy;
      hz = cz;
    }
    double *mx = C->harena.alloc<double>(lv.g.nx), *my = C->harena.alloc<double>(lv.g.ny),
           *mz = C->harena.alloc<double>(lv.g.nz);
    HIPCHK(hipMemcpyAsync(mx, hx.data(), hx.size() * 8, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(my, hy.data(), hy.size() * 8, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(mz, hz.data(), hz.size() * 8, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));  // hx, hy, hz change in the next pass
    H.M[l] = Metric{mx, my, mz};
    lv.r = C->harena.alloc<double>(N);     // f of this level
    lv.e[0] = C->harena.alloc<double>(N);  // phi
    if (l == 0) lv.e[1] = C->harena.alloc<double>(N);  // phi of the previous cycle
    for (int q = 0; q < 4; q++) H.K[l][q] = C->harena.alloc<double>(N);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(stream));
  H.built = true;
}

// Linear::solve of one level by CG with the diagonal preconditioner
// (call_smoother / call_solver): phi is improved in place
template <bool SYM>
void ref_cg(psb_ctx *C, int l, int maxit, double res_rat, double res_tol, double scale,
            int stalecount) {
  RefHier &H = C->href;
  const Level<double> &lv = H.L[l];
  const Grid &g = lv.g;
  double *r = H.K[l][0], *p0 = H.K[l][1], *p1 = H.K[l][2], *q = H.K[l][3];
  DevState *st = C->st64;
  hipLaunchKernelGGL(k_stinit, dim3(1), dim3(64), 0, C->stream, st, maxit, res_rat, res_tol,
                     scale, (double)g.N, stalecount);
  zero(C, p0, g.N);
  zero(C, p1, g.N);
  start<double, SYM, 0>(C, g, lv.mat(), lv.r, lv.e[0], r, nullptr, 0.0, st, &H.M[l]);
  for (int it = 0; it < maxit; it++) {
    double *pold = (it & 1) ? p1 : p0, *pnew = (it & 1) ? p0 : p1;
    cg_pq<double, SYM, true>(C, g, lv.mat(), r, pold, pnew, q, nullptr, st);
    cg_xr<double, true>(C, g, lv.c, pnew, q, lv.e[0], r, nullptr, 0.0, st, -1, &H.M[l]);
  }
}

// Linear::solve of one level by BiCGS with the diagonal preconditioner
// (bicgs.cpp:9-229): AC::call_smoother / call_solver hand every level the
// equation's own Krylov solver (additive_cycle_functions.cpp:89-107), which
// for the non-symmetric equations is BiCGS.  The kernels and scalar steps
// of the single-level BiCGS solve (solve_bicgs_diag) on the level's arrays;
// phi is improved in place.  Work vectors: the CG's four and four more
// (ref_bicgs_buffers, before any capture).
void ref_bicgs_buffers(psb_ctx *C) {
  RefHier &H = C->href;
  for (size_t l = 0; l < H.L.size(); l++)
    while (H.K[l].size() < 8) H.K[l].push_back(C->harena.alloc<double>(H.L[l].g.N));
}

template <bool SYM>
void ref_bicgs(psb_ctx *C, int l, int maxit, double res_rat, double res_tol, double scale,
               int stalecount) {
  RefHier &H = C->href;
  const Level<double> &lv = H.L[l];
  const Grid &g = lv.g;
  if (H.K[l].size() < 8) throw HipError{hipErrorInvalidValue};  // buffers not made
  double *r = H.K[l][0], *rt = H.K[l][4], *s = H.K[l][5], *t = H.K[l][6];
  double *p[2] = {H.K[l][1], H.K[l][2]}, *v[2] = {H.K[l][3], H.K[l][7]};
  DevState *st = C->st64;
  const DMat<double> A = lv.mat();
  hipLaunchKernelGGL(k_stinit, dim3(1), dim3(64), 0, C->stream, st, maxit, res_rat, res_tol,
                     scale, (double)g.N, stalecount);
  for (int q = 0; q < 2; q++) {
    zero(C, p[q], g.N);
    zero(C, v[q], g.N);
  }
  start<double, SYM, 2>(C, g, A, lv.r, lv.e[0], r, rt, 0.0, st, &H.M[l]);
  const int nb = nblocks<double>(g);
  for (int it = 0; it < maxit; it++) {
    const int par = it & 1;
    with_v<double>(g, [&](auto vt) {
      constexpr int V = decltype(vt)::value;
      hipLaunchKernelGGL((k_bi1<double, V, SYM>), dim3(nb), dim3(kBlock), 0, C->stream, g, A, r,
                         rt, p[par], v[par], p[par ^ 1], v[par ^ 1], st, C->work);
      finalize<1>(C, FIN_BI1, nb, st);
      hipLaunchKernelGGL((k_bi2<double, V, SYM>), dim3(nb), dim3(kBlock), 0, C->stream, g, A, r,
                         v[par ^ 1], s, t, st, C->work);
      finalize<3>(C, FIN_BI2, nb, st);
      hipLaunchKernelGGL((k_bi3<double, V>), dim3(nb), dim3(kBlock), 0, C->stream, g, A.c,
                         H.M[l], p[par ^ 1], s, t, rt, lv.e[0], r, st, C->work);
      finalize<3>(C, FIN_BI3, nb, st);
    });
  }
}

// the level-solve parameters of one cycle: MaxIter and stalecount per stage
// (pre, re-smoothing, post), AC::resrat / restol, the smoothers' scale 1/dt,
// the levels' Krylov solver (PSB_CG or PSB_BICGS)
struct RefPar {
  int mv[3], ms[3];
  double rr, rt, scale;
  int solver;
};

void ref_solve(psb_ctx *C, int l, int maxit, double rr, double rt, double scale, int stale,
               const RefPar &q) {
  // levels of a non-symmetric matrix hold 7 arrays (build_ref)
  if (q.solver == PSB_BICGS) {
    if (C->sym) ref_bicgs<true>(C, l, maxit, rr, rt, scale, stale);
    else ref_bicgs<false>(C, l, maxit, rr, rt, scale, stale);
  } else {
    if (C->sym) ref_cg<true>(C, l, maxit, rr, rt, scale, stale);
    else ref_cg<false>(C, l, maxit, rr, rt, scale, stale);
  }
}

void ref_interp(psb_ctx *C, int lc) {  // level lc into lc - 1 (h.phi += H.phi)
  RefHier &H = C->href;
  const Level<double> &cL = H.L[lc], &f = H.L[lc - 1];
  hipLaunchKernelGGL(k_ac_interp, dim3(grid_blocks(f.g.N)), dim3(kBlock), 0, C->stream, f.g,
                     f.e[0], cL.g, cL.e[0], cL.CI, cL.CJ, cL.CK);
}

void ref_restrict(psb_ctx *C, int l) {  // residual of level l into l + 1, mean removed
  RefHier &H = C->href;
  const Level<double> &f = H.L[l], &cL = H.L[l + 1];
  const int nb = grid_blocks(cL.g.N);
  hipLaunchKernelGGL(k_ac_restrict, dim3(nb), dim3(kBlock), 0, C->stream, f.g, f.mat(), f.e[0],
                     f.r, cL.g, cL.CI, cL.CJ, cL.CK, cL.aflag, cL.r, C->work);
  finalize<2>(C, FIN_ACMEAN, nb, C->st64);
  hipLaunchKernelGGL(k_ac_submean, dim3(nb), dim3(kBlock), 0, C->stream, cL.g, cL.r, cL.aflag,
                     C->st64);
}

void ref_coarsest(psb_ctx *C, int l, const RefPar &q) {  // MaxIter 100, ResRat -1
  ref_solve(C, l, 100, -1.0, q.rt, q.scale, -1, q);
  ref_interp(C, l);
}

// v_cycle (additive_recursive.cpp:4-34)
void ref_v(psb_ctx *C, int l, const RefPar &q) {
  if (l != (int)C->href.L.size() - 1) {
    ref_solve(C, l, q.mv[0], q.rr, q.rt, q.scale, q.ms[0], q);
    ref_restrict(C, l);
    ref_v(C, l + 1, q);
    ref_solve(C, l, q.mv[2], q.rr, q.rt, q.scale, q.ms[2], q);
    if (l > 0) ref_interp(C, l);
  } else {
    ref_coarsest(C, l, q);
  }
}

// f_cycle (:37-79) and w_cycle (:82-124): after the first descent (an F-
// or a W-cycle) the level is smoothed again (MaxIter mv[1]), its residual
// restricted once more and a second descent made (F: a V-cycle, W: a
// W-cycle) on the coarse phi the first one left
void ref_fw(psb_ctx *C, int l, const RefPar &q, bool w) {
  if (l != (int)C->href.L.size() - 1) {
    ref_solve(C, l, q.mv[0], q.rr, q.rt, q.scale, q.ms[0], q);
    ref_restrict(C, l);
    ref_fw(C, l + 1, q, w);
    ref_solve(C, l, q.mv[1], q.rr, q.rt, q.scale, q.ms[1], q);
    ref_restrict(C, l);
    if (w) ref_fw(C, l + 1, q, true);
    else ref_v(C, l + 1, q);
    ref_solve(C, l, q.mv[2], q.rr, q.rt, q.scale, q.ms[2], q);
    if (l > 0) ref_interp(C, l);
  } else {
    ref_coarsest(C, l, q);
  }
}

void ref_loop(psb_ctx *C, int l, int shape, const RefPar &q) {
  if (shape == PSB_CYCLE_V) ref_v(C, l, q);
  else ref_fw(C, l, q, shape == PSB_CYCLE_W);
}

// flex_cycle (additive_recursive.cpp:161-200) from level 0, its tail
// recursion as a loop driven from the host (each step depends on whether
// the level's CG met its tolerance): smooth (MaxIter mv[0] going down,
// mv[2] coming up); a level whose smoother converged hands its phi up
// (interpolation) and the finer level smooths again, level 0 ends the
// cycle; one whose smoother did not restricts its residual and goes down;
// the coarsest level is solved and hands up.  Returns false after
// max_visits level visits (the reference does not bound them).
bool ref_flex(psb_ctx *C, const RefPar &q, int max_visits) {
  const int nl = (int)C->href.L.size();
  int l = 0;
  bool up = false;
  for (int v = 0; v < max_visits; v++) {
    if (l != nl - 1) {
      const int s = up ? 2 : 0;
      ref_solve(C, l, q.mv[s], q.rr, q.rt, q.scale, q.ms[s], q);
      read_state(C, C->st64);
      if (C->hst->converged) {
        if (l == 0) return true;
        ref_interp(C, l);
        l--;
        up = true;
      } else {
        ref_restrict(C, l);
        l++;
        up = false;
      }
    } else {
      ref_coarsest(C, l, q);
      l--;
      up = false;  // the call's default sign (additive.h:185-187)
    }
  }
  return false;
}

// full_cycle (:127-163): restrict down to the coarsest level, then the
// chosen cycle from every level on the way back
void ref_full(psb_ctx *C, int l, int shape, const RefPar &q) {
  if (l != (int)C->href.L.size() - 1) {
    ref_restrict(C, l);
    ref_full(C, l + 1, shape, q);
    ref_loop(C, l, shape, q);
  } else {
    ref_coarsest(C, l, q);
  }
}

// AC::residual (additive_residual.cpp:19-29) of level 0: sqrt(dot_voldiv_avg
// (f - A phi)) * dt, and max |f - A phi| / dV * dt (Scalar::max_abs_voldiv,
// scalar_minmax.cpp:155-165) into *linf
double ref_residual(psb_ctx *C, double *linf) {
  RefHier &H = C->href;
  const Level<double> &lv = H.L[0];
  const Grid &g = lv.g;
  auto go = [&](auto symtag) {
    constexpr bool S = decltype(symtag)::value;
    with_v<double>(g, [&](auto vt) {
      constexpr int V = decltype(vt)::value;
      hipLaunchKernelGGL((k_start<double, V, S, 0>), dim3(nblocks<double>(g)), dim3(kBlock), 0,
                         C->stream, g, lv.mat(), H.M[0], lv.r, lv.e[0], H.K[0][0], nullptr,
                         0.0, C->work);
    });
  };
  if (C->sym) go(std::true_type());
  else go(std::false_type());
  finalize<3>(C, FIN_ACRES, nblocks<double>(g), C->st32, nullptr, C->relres);
  const int nb = grid_blocks(g.N);
  hipLaunchKernelGGL(k_absmax_voldiv, dim3(nb), dim3(kBlock), 0, C->stream, g, H.M[0],
                     H.K[0][0], C->work.partial);
  hipLaunchKernelGGL(k_max_final, dim3(1), dim3(kBlock), 0, C->stream, C->work.partial, nb,
                     C->relres + 1);
  double v[2] = {0.0, 0.0};
  HIPCHK(hipMemcpyAsync(v, C->relres, 16, hipMemcpyDeviceToHost, C->stream));
  HIPCHK(hipStreamSynchronize(C->stream));
  *linf = v[1] * C->ref_dt;
  return v[0];
}

// AC::cycle (additive_cycle.cpp:4-99) with init_cycles and converged
// (additive_cycle_functions.cpp:12-87).  The cycles run as captured graphs
// (one for the loop's cycle, one for the full-cycle start); the stop tests
// read the residual on the host after each, as the reference does.
int ac_cycle(psb_ctx *C, double *x, const double *b, const psb_cycle *P, psb_result *R) {
  if (P->max_cycles < 0 || P->min_cycles < 0 || P->dt <= 0.0 || P->max_iter[0] < 1 ||
      P->max_iter[1] < 1 || P->max_iter[2] < 1)
    return PSB_EINVAL;
  if (P->loop < PSB_CYCLE_V || P->loop > PSB_CYCLE_FLEX) return PSB_EINVAL;
  if (P->loop == PSB_CYCLE_FLEX && P->max_visits < 1) return PSB_EINVAL;
  if (P->init < PSB_CYCLE_NONE || P->init > PSB_CYCLE_W) return PSB_EINVAL;
  if (P->level_solver != PSB_CG && P->level_solver != PSB_BICGS) return PSB_EINVAL;
  for (int s = 0; s < 3; s++)
    if (P->max_stale[s] > kStaleMax) return PSB_EINVAL;
  build_ref(C);
  if (P->level_solver == PSB_BICGS) ref_bicgs_buffers(C);
  RefHier &H = C->href;
  const Level<double> &L0 = H.L[0];
  const size_t N0 = L0.g.N;
  HIPCHK(hipMemcpyAsync(L0.e[0], x, N0 * 8, hipMemcpyDeviceToDevice, C->stream));
  HIPCHK(hipMemcpyAsync(L0.r, b, N0 * 8, hipMemcpyDeviceToDevice, C->stream));
  // the residual norm's state: N of level 0 and dt
  psb_stop s;
  psb_stop_defaults(&s);
  s.scale = P->dt;  // AC::residual multiplies by dt
  C->ref_dt = P->dt;
  init_state(C, C->st32, &s, 0, 0.0, 0.0, 1, false);
  const double toler = P->res_tol, factor = P->res_rat, dti = 1.0 / P->dt;
  const bool linf = P->use_linf != 0;
  C->hist_host.clear();
  double linf_0 = 0.0;
  const double res_0 = ref_residual(C, &linf_0);
  C->hist_host.push_back(linf ? linf_0 : res_0);
  // fold: the last "good" phi (L0.e[1])
  HIPCHK(hipMemcpyAsync(L0.e[1], L0.e[0], N0 * 8, hipMemcpyDeviceToDevice, C->stream));
  int ncyc = 0, conv = 0, status = PSB_MAXITER;
  double res_out = linf ? linf_0 : res_0;
  const double r0c = linf ? linf_0 : res_0;
  if ((r0c < toler && P->min_cycles == 0) || res_0 == 0.0) {
    conv = 1;
    status = PSB_OK;
  } else {
    for (int l = 1; l < (int)H.L.size(); l++) {
      zero(C, H.L[l].e[0], H.L[l].g.N);
      zero(C, H.L[l].r, H.L[l].g.N);
    }
    const RefPar q{{P->max_iter[0], P->max_iter[1], P->max_iter[2]},
                   {P->max_stale[0], P->max_stale[1], P->max_stale[2]}, P->smooth_res_rat,
                   P->smooth_res_tol, dti, P->level_solver};
    // the graphs hold the cycle's shape, MaxIter values and level solver; the
    // tolerances and the scale are kernel arguments compared separately
    const long long key = (long long)q.mv[0] + ((long long)q.mv[1] << 16) +
                          ((long long)q.mv[2] << 32) + ((long long)P->loop << 48) +
                          ((long long)P->level_solver << 60);
    const long long fkey = key + ((long long)(P->init + 1) << 52);
    const bool same_ms = std::equal(q.ms, q.ms + 3, H.gms);
    if (H.gr.key != key || H.gr.fkey != dti || H.grr != q.rr || H.grt != q.rt || !same_ms) {
      destroy_graph(H.gr);
      destroy_graph(H.gfull);
    }
    H.gr.fkey = dti;
    H.grr = q.rr;
    H.grt = q.rt;
    std::copy(q.ms, q.ms + 3, H.gms);
    // AC::converged: 0 converged, 1 diverging (fold restored), 2 go on
    auto converged = [&](int cycle, double &res1, double &linf1, double res0,
                         double linf0) {
      res1 = ref_residual(C, &linf1);
      const double rc = linf ? linf1 : res1;
      const double prev = linf ? linf0 : res0;
      const double beg = linf ? linf_0 : res_0;
      if (cycle >= P->min_cycles && (rc / beg < factor || rc < toler)) return 0;
      bool check = P->stop_if_div != 0;
      if (P->priority_min_cyc && cycle < P->min_cycles) check = false;
      if (check && rc >= prev) {
        HIPCHK(hipMemcpyAsync(L0.e[0], L0.e[1], N0 * 8, hipMemcpyDeviceToDevice, C->stream));
        return 1;
      }
      return 2;
    };
    double res_m0 = 0.0, linf_m0 = 0.0, res_m1 = 0.0, linf_m1 = 0.0;
    if (P->init != PSB_CYCLE_NONE) {
      capture(C, H.gfull, fkey, [&]() { ref_full(C, 0, P->init, q); });
      HIPCHK(hipGraphLaunch(H.gfull.exec, C->stream));
      res_m1 = ref_residual(C, &linf_m1);
      // its verdict is not used (additive_cycle.cpp:40-42); with the
      // divergence check on, the equal residuals restore fold
      (void)converged(0, res_m0, linf_m0, res_m1, linf_m1);
      C->hist_host.push_back(linf ? linf_m0 : res_m0);
      res_out = linf ? linf_m0 : res_m0;
    } else {
      res_m1 = ref_residual(C, &linf_m1);
    }
    for (int c = 1; c <= P->max_cycles; c++) {
      if (P->loop == PSB_CYCLE_FLEX) {
        if (!ref_flex(C, q, P->max_visits)) {
          // no end in sight: the last cycle's phi, as after a divergence
          HIPCHK(hipMemcpyAsync(L0.e[0], L0.e[1], N0 * 8, hipMemcpyDeviceToDevice, C->stream));
          ncyc = c;
          status = PSB_DIVERGED;
          break;
        }
      } else {
        capture(C, H.gr, key, [&]() { ref_loop(C, 0, P->loop, q); });
        HIPCHK(hipGraphLaunch(H.gr.exec, C->stream));
      }
      const int con = converged(c, res_m0, linf_m0, res_m1, linf_m1);
      C->hist_host.push_back(linf ? linf_m0 : res_m0);
      res_out = linf ? linf_m0 : res_m0;
      res_m1 = res_m0;  // linf_m1 keeps its value (additive_cycle.cpp:78)
      ncyc = c;
      if (con == 0) {
        conv = 1;
        status = PSB_OK;
        break;
      }
      if (con == 1) {
        status = PSB_DIVERGED;
        break;
      }
      HIPCHK(hipMemcpyAsync(L0.e[1], L0.e[0], N0 * 8, hipMemcpyDeviceToDevice, C->stream));
    }
  }
  HIPCHK(hipMemcpyAsync(x, L0.e[0], N0 * 8, hipMemcpyDeviceToDevice, C->stream));
  HIPCHK(hipStreamSynchronize(C->stream));
  C->hist_on_host = true;
  C->hist_count = (int)C->hist_host.size();
  R->iterations = ncyc;
  R->outer = 0;
  R->converged = conv;
  R->status = status;
  R->res0 = r0c;
  R->res = res_out;
  return status;
}

int solve_device(psb_ctx *C, double *x, const double *b, const psb_stop *p, psb_result *R) {
  C->hist_on_host = false;
  if (p->precond == PSB_PREC_AC) {
    if (p->method == PSB_BICGS) {  // the V-cycle as BiCGS's preconditioner, FP64
      if ((p->precision != PSB_FP64 && !(p->precision == PSB_MIXED && C->sym)) ||
          p->ac_pre < 1 || p->ac_post < 1 || p->ac_coarsen < 0 || p->ac_coarsen > 1)
        return PSB_EINVAL;
      ensure_rule(C, p->ac_coarsen);
      C->cg_half = false;
      C->fuse = false;
      C->kmask = 0;
      C->flex = false;
      C->ac_omega = p->ac_omega;
      if (p->precision == PSB_MIXED) solve_acbicgs_mixed(C, x, b, p, R);
      else if (C->sym) solve_acbicgs64<true>(C, x, b, p, R);
      else solve_acbicgs64<false>(C, x, b, p, R);
      R->rel_l2 = true_rel(C, b, x);
      return R->status;
    }
    if (p->method != PSB_CG || !C->sym || p->ac_pre < 1 || p->ac_post < 1) return PSB_EINVAL;
    if (p->ac_coarsen < 0 || p->ac_coarsen > 1) return PSB_EINVAL;
    if (p->ac_store < 0 || p->ac_store > 2) return PSB_EINVAL;
    if (p->ac_kcycle < -1 || (p->ac_kcycle > 0 && (p->ac_kcycle & 1))) return PSB_EINVAL;
    if (p->ac_flexible < -1 || p->ac_flexible > 1) return PSB_EINVAL;
    ensure_rule(C, p->ac_coarsen);
    // settings of this solve, read while its graph is captured (the graph
    // cache keys on them: no setting change drops a graph)
    C->cg_half = p->precision == PSB_MIXED && p->ac_store == 2;
    C->fuse = p->ac_fuse != 0;
    C->kmask = p->ac_kcycle;
    C->flex_req = p->ac_flexible;  // resolved against the hierarchy (set_flex)
    C->ac_omega = p->ac_omega;
    if (p->precision == PSB_MIXED) {
      build_hier32(C);
      set_half(C, p->ac_store >= 1);
    }
    if (p->precision == PSB_MIXED && p->inner_adapt == 2) {
      solve_ru(C, x, b, p, R);  // sets rel_l2 itself
      // x is final when the call returns (its last update may be in flight)
      HIPCHK(hipStreamSynchronize(C->stream));
      return R->status;
    } else if (p->precision == PSB_MIXED)
      solve_ir(C, x, b, p, R);
    else
      solve_ac64(C, x, b, p, R);
  } else if (p->precond == PSB_PREC_SGS) {
    if (p->precision != PSB_FP64 || p->method != PSB_CG) return PSB_EINVAL;
    if (C->sym) solve_cg_sgs<true>(C, x, b, p, R);
    else solve_cg_sgs<false>(C, x, b, p, R);
  } else {
    if (p->precond != PSB_PREC_DIAG || p->precision != PSB_FP64) return PSB_EINVAL;
    if (p->method == PSB_CG) {
      if (C->sym) solve_cg_diag<true>(C, x, b, p, R);
      else solve_cg_diag<false>(C, x, b, p, R);
    } else if (p->method == PSB_BICGS) {
      if (C->sym) solve_bicgs_diag<true>(C, x, b, p, R);
      else solve_bicgs_diag<false>(C, x, b, p, R);
    } else {
      return PSB_EINVAL;
    }
  }
  R->rel_l2 = true_rel(C, b, x);
  return R->status;
}

// host layout (optional ghost padding) <-> compact
size_t host_index(const psb_layout &L, int i, int j, int k) {
  const int G = L.ghost;
  const size_t ny = L.ny + 2 * G, nz = L.nz + 2 * G;
  return ((size_t)(i + G) * ny + (j + G)) * nz + (k + G);
}

void gather_host(const psb_layout &L, const double *src, std::vector<double> &dst) {
  dst.resize((size_t)L.nx * L.ny * L.nz);
  size_t q = 0;
  for (int i = 0; i < L.nx; i++)
    for (int j = 0; j < L.ny; j++) {
      const double *row = src + host_index(L, i, j, 0);
      std::memcpy(&dst[q], row, sizeof(double) * L.nz);
      q += L.nz;
    }
}

void scatter_host(const psb_layout &L, const std::vector<double> &src, double *dst) {
  size_t q = 0;
  for (int i = 0; i < L.nx; i++)
    for (int j = 0; j < L.ny; j++) {
      std::memcpy(dst + host_index(L, i, j, 0), &src[q], sizeof(double) * L.nz);
      q += L.nz;
    }
}


}  // namespace host
}  // namespace psb

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

void psb_stop_defaults(psb_stop *s) {
  std::memset(s, 0, sizeof(*s));
  s->struct_size = (int)sizeof(psb_stop);
  s->min_iter = 1;
  s->max_iter = 20;
  s->res_rat = 0.01;
  s->res_tol = 1.0e-15;
  s->scale = 1.0;
  s->stalecount = -1;
  s->precform = 1;
  s->method = PSB_CG;
  s->precond = PSB_PREC_DIAG;
  s->precision = PSB_FP64;
  s->norm = PSB_NORM_REF;
  s->ac_pre = 2;
  s->ac_post = 2;
  s->ac_omega = 0.9;
  s->ac_cfac = 2.0;
  s->inner_rtol = 1.0e-4;
  s->max_outer = 30;
  s->graph_chunk = 8;
  s->inner_adapt = 0;
  s->ac_coarsen = 0;
  s->ac_store = 0;
  s->ac_fuse = 0;
  s->ac_kcycle = 0;
  s->ac_flexible = -1;
}

const char *psb_status_string(int s) {
  switch (s) {
    case PSB_OK: return "converged";
    case PSB_MAXITER: return "maximum iterations reached";
    case PSB_STALE: return "stalled";
    case PSB_BREAKDOWN: return "breakdown";
    case PSB_DIVERGED: return "diverged";
    case PSB_NAN: return "non-finite residual";
    case PSB_EINVAL: return "invalid argument";
    case PSB_EHIP: return "HIP runtime error";
    case PSB_ENOMEM: return "out of device memory";
    case PSB_ENOTREADY: return "matrix not set";
    default: return "unknown status";
  }
}

int psb_create(psb_ctx **out, const psb_layout *L) {
  if (!out || !L) return PSB_EINVAL;
  *out = nullptr;
  if (L->nx < 1 || L->ny < 1 || L->nz < 1 || L->ghost < 0) return PSB_EINVAL;
  if ((long long)L->nx * L->ny * L->nz >= (1LL << 31)) return PSB_EINVAL;
  psb_ctx *C = new (std::nothrow) psb_ctx();
  if (!C) return PSB_ENOMEM;
  int rc = guarded(C, [&]() {
    C->lay = *L;
    C->g = make_grid(L->nx, L->ny, L->nz);
    // a blocking stream: work the caller queued on the null stream (a
    // hipMemset of x, say) completes before the context's kernels read it
    HIPCHK(hipStreamCreateWithFlags(&C->stream, hipStreamDefault));
    C->st64 = C->base.alloc<DevState>(1);
    C->st32 = C->base.alloc<DevState>(1);
    HIPCHK(hipHostMalloc((void **)&C->hst, sizeof(DevState), hipHostMallocDefault));
    std::memset(C->hst, 0, sizeof(DevState));
    C->work.partial = C->base.alloc<double>((size_t)kMaxBlocks * kMaxPartials);
    C->relres = C->base.alloc<double>(2);
    const int N = C->g.N;
    C->dx = C->base.alloc<double>(N);
    C->db = C->base.alloc<double>(N);
    C->dy = C->base.alloc<double>(N);
    C->static_scratch = static_scratch_here();
    // which fine-level kernels use the LDS-staged form (same results):
    // PSB_NO_LDS set: none (for comparisons); PSB_LDS_MASK="m32,m64"
    // overrides the measured defaults (bits kLdsSweep .. kLdsCgPq)
    if (std::getenv("PSB_NO_KFUSE")) C->kfuse = false;
    if (std::getenv("PSB_NO_LDS")) {
      C->lds32 = C->lds64 = 0;
    } else if (const char *m = std::getenv("PSB_LDS_MASK")) {
      if (std::sscanf(m, "%d,%d", &C->lds32, &C->lds64) != 2)
        throw std::invalid_argument("PSB_LDS_MASK is not \"m32,m64\"");
    }
    return PSB_OK;
  });
  if (rc != PSB_OK) {
    psb_destroy(C);
    return rc;
  }
  *out = C;
  return PSB_OK;
}

int psb_destroy(psb_ctx *C) {
  if (!C) return PSB_OK;
  if (C->stream) (void)hipStreamSynchronize(C->stream);
  destroy_graphs(C);
  C->arena.free_all();
  C->harena.free_all();
  C->base.free_all();
  if (C->hst) (void)hipHostFree(C->hst);
  if (C->stream) (void)hipStreamDestroy(C->stream);
  delete C;
  return PSB_OK;
}

int psb_set_geometry(psb_ctx *C, const double *dxc, const double *dyc, const double *dzc) {
  if (!C) return PSB_EINVAL;
  if ((dxc == nullptr) != (dyc == nullptr) || (dxc == nullptr) != (dzc == nullptr))
    return PSB_EINVAL;
  return guarded(C, [&]() {
    HIPCHK(hipStreamSynchronize(C->stream));
    destroy_graphs(C);
    C->version++;
    reset_hier(C);  // the reference cycle's coarse metrics derive from these
    if (!dxc) {
      C->dxc = C->dyc = C->dzc = nullptr;
      C->gdx.clear();
      C->gdy.clear();
      C->gdz.clear();
      return PSB_OK;
    }
    C->gdx.assign(dxc, dxc + C->g.nx);
    C->gdy.assign(dyc, dyc + C->g.ny);
    C->gdz.assign(dzc, dzc + C->g.nz);
    if (!C->dxc) {
      C->dxc = C->base.alloc<double>(C->g.nx);
      C->dyc = C->base.alloc<double>(C->g.ny);
      C->dzc = C->base.alloc<double>(C->g.nz);
    }
    HIPCHK(hipMemcpy(C->dxc, dxc, C->g.nx * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(C->dyc, dyc, C->g.ny * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(C->dzc, dzc, C->g.nz * 8, hipMemcpyHostToDevice));
    return PSB_OK;
  });
}

int psb_set_matrix(psb_ctx *C, const psb_matrix7 *A) {
  if (!C || !A || !A->c || !A->w || !A->e || !A->s || !A->n || !A->b || !A->t) return PSB_EINVAL;
  return guarded(C, [&]() {
    const psb_layout &L = C->lay;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    std::vector<double> a[7];
    const double *src[7] = {A->c, A->w, A->s, A->b, A->e, A->n, A->t};
    for (int q = 0; q < 7; q++) gather_host(L, src[q], a[q]);
    const std::vector<double> &w = a[1], &s = a[2], &b = a[3], &e = a[4], &n = a[5], &t = a[6];
    auto id = [&](int i, int j, int k) { return ((size_t)i * ny + j) * nz + k; };
    // wall couplings must already be folded into the system (zero); the
    // operator is singular when every active row sums to zero (Neumann
    // walls, no time term)
    bool sym = true, singular = true;
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          size_t P = id(i, j, k);
          if ((i == 0 && w[P] != 0.0) || (i == nx - 1 && e[P] != 0.0) ||
              (j == 0 && s[P] != 0.0) || (j == ny - 1 && n[P] != 0.0) ||
              (k == 0 && b[P] != 0.0) || (k == nz - 1 && t[P] != 0.0))
            return PSB_EINVAL;
          if (i < nx - 1 && e[P] != w[id(i + 1, j, k)]) sym = false;
          if (j < ny - 1 && n[P] != s[id(i, j + 1, k)]) sym = false;
          if (k < nz - 1 && t[P] != b[id(i, j, k + 1)]) sym = false;
          const double cP = a[0][P], off = w[P] + e[P] + s[P] + n[P] + b[P] + t[P];
          const bool active = C->off.empty() || !C->off[P];
          if (active && std::fabs(cP - off) > 1e-10 * std::fabs(cP)) singular = false;
        }
    const int N = C->g.N;
    const int nstore = sym ? 4 : 7;
    // (re)allocate: a new matrix invalidates the hierarchy and graphs
    HIPCHK(hipStreamSynchronize(C->stream));
    destroy_graphs(C);
    C->arena.release();
    C->harena.release();
    for (int q = 0; q < 7; q++) C->A64[q] = nullptr;
    C->A32[0] = C->A32[1] = C->A32[2] = C->A32[3] = nullptr;
    C->have32 = false;
    C->h64 = Hier<double>();
    C->h32 = Hier<float>();
    C->href = RefHier();
    C->k64 = KrylovVecs<double>();
    C->k32 = KrylovVecs<float>();
    C->hist_count = 0;
    for (int q = 0; q < nstore; q++) {
      C->A64[q] = C->arena.alloc<double>(N);
      HIPCHK(hipMemcpyAsync(C->A64[q], a[q].data(), (size_t)N * 8, hipMemcpyHostToDevice,
                            C->stream));
    }
    C->aflag0 = nullptr;
    C->metrics = DevMetrics{};
    C->off_dev = nullptr;
    if (!C->off.empty()) {
      std::vector<uint8_t> fl(N);
      for (int q = 0; q < N; q++) fl[q] = C->off[q] ? 0 : 1;
      C->aflag0 = C->arena.alloc<uint8_t>(N);
      HIPCHK(hipMemcpyAsync(C->aflag0, fl.data(), N, hipMemcpyHostToDevice, C->stream));
    }
    HIPCHK(hipStreamSynchronize(C->stream));
    C->sym = sym;
    C->singular = singular;
    C->have_matrix = true;
    C->version++;
    return PSB_OK;
  });
}

namespace {

// drop everything built from the previous matrix (graphs, hierarchies,
// Krylov vectors, the matrix itself)
void reset_matrix(psb_ctx *C) {
  HIPCHK(hipStreamSynchronize(C->stream));
  destroy_graphs(C);
  C->arena.release();
  C->harena.release();
  for (int q = 0; q < 7; q++) C->A64[q] = nullptr;
  C->fbnd = nullptr;
  C->A32[0] = C->A32[1] = C->A32[2] = C->A32[3] = nullptr;
  C->have32 = false;
  C->h64 = Hier<double>();
  C->h32 = Hier<float>();
  C->href = RefHier();
  C->k64 = KrylovVecs<double>();
  C->k32 = KrylovVecs<float>();
  C->hist_count = 0;
  C->have_matrix = false;
}

int assemble_pressure(psb_ctx *C, const double *rho, bool rho_dev, const psb_metrics *m,
                      const unsigned char *off) {
  const double *arrs[9] = {m->dxc, m->dxw, m->dxe, m->dyc, m->dys, m->dyn, m->dzc, m->dzb, m->dzt};
  for (const double *a : arrs)
    if (!a) return PSB_EINVAL;
  const Grid &g = C->g;
  const int len[9] = {g.nx, g.nx, g.nx, g.ny, g.ny, g.ny, g.nz, g.nz, g.nz};
  int rc = psb_set_geometry(C, m->dxc, m->dyc, m->dzc);
  if (rc != PSB_OK) return rc;
  reset_matrix(C);
  double *dm[9];
  for (int q = 0; q < 9; q++) {
    dm[q] = C->arena.alloc<double>(len[q]);
    HIPCHK(hipMemcpyAsync(dm[q], arrs[q], len[q] * 8, hipMemcpyHostToDevice, C->stream));
  }
  const DevMetrics M{dm[0], dm[1], dm[2], dm[3], dm[4], dm[5], dm[6], dm[7], dm[8]};
  const size_t NG = (size_t)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
  const double *drho = rho;
  if (!rho_dev) {
    double *t = C->arena.alloc<double>(NG);
    HIPCHK(hipMemcpyAsync(t, rho, NG * 8, hipMemcpyHostToDevice, C->stream));
    drho = t;
  }
  uint8_t *doff = nullptr;
  if (off) {
    C->off.assign(off, off + g.N);
    doff = C->arena.alloc<uint8_t>(g.N);
    HIPCHK(hipMemcpyAsync(doff, C->off.data(), g.N, hipMemcpyHostToDevice, C->stream));
    std::vector<uint8_t> fl(g.N);
    for (int q = 0; q < g.N; q++) fl[q] = C->off[q] ? 0 : 1;
    C->aflag0 = C->arena.alloc<uint8_t>(g.N);
    HIPCHK(hipMemcpyAsync(C->aflag0, fl.data(), g.N, hipMemcpyHostToDevice, C->stream));
  } else {
    C->off.clear();
    C->aflag0 = nullptr;
  }
  for (int q = 0; q < 4; q++) C->A64[q] = C->arena.alloc<double>(g.N);
  // the boundary source exists when a side is Dirichlet or has a value
  bool src = false;
  for (int q = 0; q < 6; q++) src = src || C->bnd.type[q] != PSB_BND_NEUMANN || C->bnd.value[q] != 0.0;
  C->fbnd = src ? C->arena.alloc<double>(g.N) : nullptr;
  hipLaunchKernelGGL(k_discretize, dim3(grid_blocks(g.N)), dim3(kBlock), 0, C->stream, g, M,
                     drho, doff, C->bnd, C->A64[0], C->A64[1], C->A64[2], C->A64[3], C->fbnd);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(C->stream));  // host sources above are the caller's or local
  C->metrics = M;
  C->off_dev = doff;
  C->sym = true;
  C->singular = true;  // unless a side is Dirichlet
  for (int q = 0; q < 6; q++) C->singular = C->singular && C->bnd.type[q] == PSB_BND_NEUMANN;
  C->have_matrix = true;
  C->version++;
  return PSB_OK;
}

}  // namespace

int psb_assemble_pressure(psb_ctx *C, const double *rho, const psb_metrics *m,
                          const unsigned char *off) {
  if (!C || !rho || !m) return PSB_EINVAL;
  return guarded(C, [&]() { return assemble_pressure(C, rho, false, m, off); });
}

int psb_assemble_pressure_device(psb_ctx *C, const double *d_rho, const psb_metrics *m,
                                 const unsigned char *off) {
  if (!C || !d_rho || !m) return PSB_EINVAL;
  return guarded(C, [&]() { return assemble_pressure(C, d_rho, true, m, off); });
}

int psb_update_rhs_device(psb_ctx *C, const double *d_u, const double *d_v, const double *d_w,
                          int zero_mean, double *d_b) {
  if (!C || !d_u || !d_v || !d_w || !d_b) return PSB_EINVAL;
  if (!C->have_matrix || !C->metrics.dxc) return PSB_ENOTREADY;
  return guarded(C, [&]() {
    const Grid &g = C->g;
    const int nb = grid_blocks(g.N);
    hipLaunchKernelGGL(k_update_rhs, dim3(nb), dim3(kBlock), 0, C->stream, g, C->metrics, d_u,
                       d_v, d_w, C->off_dev, C->fbnd, d_b, C->work);
    if (zero_mean) {
      finalize<2>(C, FIN_ACMEAN, nb, C->st32);
      hipLaunchKernelGGL(k_rhs_center, dim3(nb), dim3(kBlock), 0, C->stream, g, d_b,
                         C->off_dev, C->st32);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(C->stream));
    return (int)PSB_OK;
  });
}

int psb_project_device(psb_ctx *C, const double *d_rho, const double *d_p, double dt, int mask,
                       double *d_u, double *d_v, double *d_w) {
  if (!C || !d_rho || !d_p || mask < 0 || mask > 7) return PSB_EINVAL;
  if (((mask & 1) && !d_u) || ((mask & 2) && !d_v) || ((mask & 4) && !d_w)) return PSB_EINVAL;
  if (!C->have_matrix || !C->metrics.dxc) return PSB_ENOTREADY;
  return guarded(C, [&]() {
    const Grid &g = C->g;
    hipLaunchKernelGGL(k_project, dim3(grid_blocks(g.N)), dim3(kBlock), 0, C->stream, g,
                       C->metrics, d_rho, C->off_dev, d_p, dt, mask, d_u, d_v, d_w);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(C->stream));
    return (int)PSB_OK;
  });
}

int psb_set_boundary(psb_ctx *C, const psb_bnd *bnd) {
  if (!C || !bnd) return PSB_EINVAL;
  for (int q = 0; q < 6; q++)
    if ((bnd->type[q] != PSB_BND_NEUMANN && bnd->type[q] != PSB_BND_DIRICHLET) ||
        !std::isfinite(bnd->value[q]))
      return PSB_EINVAL;
  C->bnd = *bnd;  // used by the next assembly
  return PSB_OK;
}

int psb_pressure_ghost_device(psb_ctx *C, const double *d_c, const double *d_kappa,
                              const double *d_rho, double sigma, double *d_fext) {
  if (!C || !d_c || !d_kappa || !d_rho || !d_fext) return PSB_EINVAL;
  if (!C->have_matrix || !C->metrics.dxc) return PSB_ENOTREADY;
  return guarded(C, [&]() {
    const Grid &g = C->g;
    hipLaunchKernelGGL(k_ghost, dim3(grid_blocks(g.N)), dim3(kBlock), 0, C->stream, g,
                       C->metrics, d_c, d_kappa, d_rho, sigma, C->off_dev, d_fext);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(C->stream));
    return (int)PSB_OK;
  });
}

int psb_save_bck_device(psb_ctx *C, const char *path, const double *d_x) {
  if (!C || !path || !d_x) return PSB_EINVAL;
  return guarded(C, [&]() {
    std::vector<double> h(C->g.N);
    HIPCHK(hipMemcpy(h.data(), d_x, h.size() * 8, hipMemcpyDeviceToHost));
    return psb_save_bck(path, C->g.nx, C->g.ny, C->g.nz, 3, h.data());
  });
}

int psb_load_bck_device(psb_ctx *C, const char *path, double *d_x) {
  if (!C || !path || !d_x) return PSB_EINVAL;
  return guarded(C, [&]() {
    std::vector<double> h(C->g.N);
    const int rc = psb_load_bck(path, C->g.nx, C->g.ny, C->g.nz, h.data(), nullptr);
    if (rc != PSB_OK) return rc;
    HIPCHK(hipMemcpy(d_x, h.data(), h.size() * 8, hipMemcpyHostToDevice));
    return (int)PSB_OK;
  });
}

int psb_set_active(psb_ctx *C, const unsigned char *off) {
  if (!C) return PSB_EINVAL;
  // takes effect at the next psb_set_matrix, which rebuilds the hierarchy
  if (!off) {
    C->off.clear();
  } else {
    const psb_layout &L = C->lay;
    C->off.resize(C->g.N);
    size_t q = 0;
    for (int i = 0; i < L.nx; i++)
      for (int j = 0; j < L.ny; j++)
        for (int k = 0; k < L.nz; k++) C->off[q++] = off[host_index(L, i, j, k)];
  }
  C->version++;
  return PSB_OK;
}

int psb_solve_device(psb_ctx *C, double *x, const double *b, const psb_stop *p, psb_result *R) {
  if (!C || !x || !b || !p || !R || !stop_ok(p)) return PSB_EINVAL;
  if (!C->have_matrix) return PSB_ENOTREADY;
  if (p->max_iter < 0 || p->scale == 0.0 || p->stalecount > kStaleMax) return PSB_EINVAL;
  std::memset(R, 0, sizeof(*R));
  return guarded(C, [&]() { return solve_device(C, x, b, p, R); });
}

int psb_solve_device_static(psb_ctx *C, double *x, const double *b, const psb_stop *p,
                            psb_result *R) {
  if (!C || !x || !b || !p || !R || !stop_ok(p)) return PSB_EINVAL;
  if (!C->have_matrix) return PSB_ENOTREADY;
  if (p->max_iter < 0 || p->scale == 0.0 || p->stalecount > kStaleMax) return PSB_EINVAL;
  if (p->method != PSB_CG || p->precond != PSB_PREC_DIAG || p->precision != PSB_FP64)
    return PSB_EINVAL;
  if ((size_t)C->g.N > kStaticMaxCells) {
    C->err = "psb_solve_device_static: " + std::to_string(C->g.N) +
             " cells exceed the static scratch (at most " + std::to_string(kStaticMaxCells) +
             " cells)";
    std::fprintf(stderr, "%s\n", C->err.c_str());
    return PSB_EINVAL;
  }
  if (!C->static_scratch) return PSB_EHIP;
  std::memset(R, 0, sizeof(*R));
  return guarded(C, [&]() {
    if (C->sym) solve_cg_diag_static<true>(C, x, b, p, R);
    else solve_cg_diag_static<false>(C, x, b, p, R);
    return R->status;
  });
}

int psb_solve(psb_ctx *C, double *x, const double *b, const psb_stop *p, psb_result *R) {
  if (!C || !x || !b || !p || !R || !stop_ok(p)) return PSB_EINVAL;
  if (!C->have_matrix) return PSB_ENOTREADY;
  return guarded(C, [&]() {
    const int N = C->g.N;
    std::vector<double> hx, hb;
    gather_host(C->lay, x, hx);
    gather_host(C->lay, b, hb);
    HIPCHK(hipMemcpyAsync(C->dx, hx.data(), (size_t)N * 8, hipMemcpyHostToDevice, C->stream));
    HIPCHK(hipMemcpyAsync(C->db, hb.data(), (size_t)N * 8, hipMemcpyHostToDevice, C->stream));
    int rc = psb_solve_device(C, C->dx, C->db, p, R);
    if (rc < 0) return rc;
    HIPCHK(hipMemcpyAsync(hx.data(), C->dx, (size_t)N * 8, hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    scatter_host(C->lay, hx, x);
    return rc;
  });
}

int psb_apply(psb_ctx *C, const double *x, const double *b, double *y) {
  if (!C || !x || !y) return PSB_EINVAL;
  if (!C->have_matrix) return PSB_ENOTREADY;
  return guarded(C, [&]() {
    const int N = C->g.N;
    double *dy = C->dy;
    std::vector<double> hx, hb, hy(N);
    gather_host(C->lay, x, hx);
    HIPCHK(hipMemcpyAsync(C->dx, hx.data(), (size_t)N * 8, hipMemcpyHostToDevice, C->stream));
    hipStream_t stream = C->stream;
    if (b) {
      gather_host(C->lay, b, hb);
      HIPCHK(hipMemcpyAsync(C->db, hb.data(), (size_t)N * 8, hipMemcpyHostToDevice, stream));
      if (C->sym) resid_k<double, true>(stream, C->g, mat64(C), C->db, C->dx, dy);
      else resid_k<double, false>(stream, C->g, mat64(C), C->db, C->dx, dy);
    } else {
      if (C->sym) spmv<double, true>(stream, C->g, mat64(C), C->dx, dy);
      else spmv<double, false>(stream, C->g, mat64(C), C->dx, dy);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hy.data(), dy, (size_t)N * 8, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    scatter_host(C->lay, hy, y);
    return PSB_OK;
  });
}

void psb_cycle_defaults(psb_cycle *c) {
  std::memset(c, 0, sizeof(*c));
  c->struct_size = (int)sizeof(psb_cycle);
  c->res_tol = -1.0;
  c->res_rat = -1.0;
  c->max_cycles = 20;  // additive.cpp:11-22
  c->min_cycles = 0;
  c->max_iter[0] = c->max_iter[1] = c->max_iter[2] = 20;
  c->dt = 1.0;
  c->loop = PSB_CYCLE_V;
  c->init = PSB_CYCLE_NONE;
  c->stop_if_div = 1;
  c->priority_min_cyc = 1;
  c->use_linf = 0;
  c->smooth_res_rat = kAtto;
  c->smooth_res_tol = kAtto;
  c->max_stale[0] = c->max_stale[1] = c->max_stale[2] = -1;  // additive.cpp:19
  c->max_visits = 1000;
  c->level_solver = PSB_CG;
}

int psb_ac_vcycle(psb_ctx *C, double *x, const double *b, const psb_cycle *P, psb_result *R) {
  if (!C || !x || !b || !P || !R || !cycle_ok(P)) return PSB_EINVAL;
  if (!C->have_matrix) return PSB_ENOTREADY;
  return guarded(C, [&]() {
    const psb_layout &L = C->lay;
    std::vector<double> hx, hb;
    gather_host(L, x, hx);
    gather_host(L, b, hb);
    const size_t N = C->g.N;
    HIPCHK(hipMemcpy(C->dx, hx.data(), N * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(C->db, hb.data(), N * 8, hipMemcpyHostToDevice));
    int rc = ac_cycle(C, C->dx, C->db, P, R);
    if (rc < 0) return rc;
    R->rel_l2 = true_rel(C, C->db, C->dx);
    HIPCHK(hipMemcpy(hx.data(), C->dx, N * 8, hipMemcpyDeviceToHost));
    scatter_host(L, hx, x);
    return rc;
  });
}

int psb_ac_vcycle_device(psb_ctx *C, double *x, const double *b, const psb_cycle *P,
                         psb_result *R) {
  if (!C || !x || !b || !P || !R || !cycle_ok(P)) return PSB_EINVAL;
  if (!C->have_matrix) return PSB_ENOTREADY;
  return guarded(C, [&]() {
    int rc = ac_cycle(C, x, b, P, R);
    if (rc >= 0) R->rel_l2 = true_rel(C, b, x);
    return rc;
  });
}

int psb_history(psb_ctx *C, double *hist, int cap) {
  if (!C || !hist || cap < 0) return PSB_EINVAL;
  if (C->hist_on_host) {
    int n = std::min(cap, (int)C->hist_host.size());
    std::copy(C->hist_host.begin(), C->hist_host.begin() + n, hist);
    return n;
  }
  if (!C->hist || C->hist_count <= 0) return 0;
  return guarded(C, [&]() {
    int n = std::min(cap, C->hist_count);
    HIPCHK(hipMemcpy(hist, C->hist, (size_t)n * 8, hipMemcpyDeviceToHost));
    return n;
  });
}

int psb_time_kernel(psb_ctx *C, int kernel, int reps, double *avg_ms, double *bytes) {
  if (!C || !avg_ms || !bytes || reps < 1) return PSB_EINVAL;
  if (!C->have_matrix || !C->sym) return PSB_ENOTREADY;
  return guarded(C, [&]() {
    hipStream_t stream = C->stream;
    const Grid &g = C->g;
    const double N = (double)g.N;
    std::function<void()> f;
    if (kernel == 0) {
      ensure_krylov(C, C->k64, false);
      double *x = C->k64.p[0], *y = C->k64.q;
      zero(C, x, g.N);
      f = [=]() { spmv<double, true>(stream, g, mat64(C), x, y); };
      *bytes = N * (4 * 8 + 2 * 8);
    } else if (kernel == 1) {
      ensure32(C);
      ensure_krylov(C, C->k32, false);
      float *x = C->k32.p[0], *y = C->k32.q;
      zero(C, x, g.N);
      f = [=]() { spmv<float, true>(stream, g, mat32(C), x, y); };
      *bytes = N * (4 * 4 + 2 * 4);
    } else if (kernel == 2) {
      build_hier32(C);
      Level<float> lv = C->h32.L[0];
      zero(C, lv.e[0], g.N);
      zero(C, lv.r, g.N);
      DevState h;
      std::memset(&h, 0, sizeof(h));
      HIPCHK(hipMemcpyAsync(C->st32, &h, sizeof(h), hipMemcpyHostToDevice, stream));
      f = [=]() {
        jacobi<float, false>(C, lv, lv.e[0], lv.e[1], nullptr, nullptr, 0.0f, 0.9f, C->st32);
      };
      // the level-0 matrix as the last mixed solve stored it (4 arrays, FP32
      // or FP16), e and r in, e out
      *bytes = N * (4 * (lv.h16() ? 2 : 4) + 3 * 4);
    } else if (kernel == 4 || kernel == 5) {
      // the two post-sweeps of the fine level: 4 in one pass (k_post2), 5 as
      // two kernels; both as the last mixed solve's hierarchy stores the level
      build_hier32(C);
      Level<float> lv = C->h32.L[0];
      const Level<float> cL = C->h32.L[1];
      zero(C, lv.e[0], g.N);
      zero(C, lv.r, g.N);
      zero(C, cL.e[0], cL.g.N);
      DevState h;
      std::memset(&h, 0, sizeof(h));
      HIPCHK(hipMemcpyAsync(C->st32, &h, sizeof(h), hipMemcpyHostToDevice, stream));
      const bool fused = kernel == 4;
      f = [=]() {
        const bool keep = C->fuse;
        C->fuse = true;  // time the fused pass whatever the last solve selected
        const bool done = fused && post2<float, false>(C, lv, lv.e[0], lv.e[2], cL, cL.e[0], 2.0f,
                                                       0.9f, 1.0f, C->st32);
        C->fuse = keep;
        if (done) return;
        jacobi<float, false>(C, lv, lv.e[0], lv.e[1], &cL, cL.e[0], 2.0f, 0.9f, C->st32);
        jacobi<float, false>(C, lv, lv.e[1], lv.e[2], nullptr, nullptr, 0.0f, 0.9f, C->st32);
      };
      *bytes = N * (4 * (lv.h16() ? 2 : 4) + 3 * 4);  // one sweep's algorithmic bytes
    } else if (kernel == 6) {
      // the FP32 CG p/q step of the mixed AC solve on level 0 (flux-form FP32
      // matrix, marching): p = z + beta p_old with beta 0, q = A p, p.q
      build_hier32(C);
      ensure_krylov(C, C->k32, false);
      const Level<float> lv = C->h32.L[0];
      const KrylovVecs<float> K = C->k32;
      MarchShape m;
      if (!lv.d || !vec_ok<float>(g) || !march_ok<4>(g, m)) return PSB_EINVAL;
      zero(C, lv.e[0], g.N);
      zero(C, K.p[0], g.N);
      DevState h;
      std::memset(&h, 0, sizeof(h));
      HIPCHK(hipMemcpyAsync(C->st32, &h, sizeof(h), hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
      const bool lds = lds_use<float>(C, kLdsCgPq) && lds_ok<4>(g);
      f = [=]() {
        if (lds)
          hipLaunchKernelGGL((k_cg_pq_lds<float, 4, true, false, true, float>), dim3(m.nblocks),
                             dim3(kBlock), 0, stream, g, lv.matflux(), lv.e[0], K.p[0], K.p[1],
                             K.q, lv.aflag, C->st32, C->work, m.TI, m.ntiles, 1.0f);
        else
          hipLaunchKernelGGL((k_cg_pq_march<float, 4, true, false, true, float>),
                             dim3(m.nblocks), dim3(kBlock), 0, stream, g, lv.matflux(), lv.e[0],
                             K.p[0], K.p[1], K.q, lv.aflag, C->st32, C->work, m.TI, m.ntiles,
                             1.0f);
      };
      // four FP32 coefficient arrays, z and p_old in, p and q out
      *bytes = N * (4 * 4 + 4 * 4 + (lv.aflag ? 1 : 0));
    } else if (kernel == 7) {
      // the first post-sweep of level 0 with the prolongation of level 1's
      // correction fused in, as the last mixed solve stores the level
      build_hier32(C);
      const Level<float> lv = C->h32.L[0];
      const Level<float> cL = C->h32.L[1];
      zero(C, lv.e[0], g.N);
      zero(C, lv.r, g.N);
      zero(C, cL.e[0], cL.g.N);
      DevState h;
      std::memset(&h, 0, sizeof(h));
      HIPCHK(hipMemcpyAsync(C->st32, &h, sizeof(h), hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
      f = [=]() {
        jacobi<float, false>(C, lv, lv.e[0], lv.e[1], &cL, cL.e[0], 2.0f, 0.9f, C->st32);
      };
      // the sweep's bytes plus one coarse value per block of fine cells (and
      // the active flags when the level has solid cells)
      *bytes = N * (4 * (lv.h16() ? 2 : 4) + 3 * 4 + 4.0 / (cL.CI * cL.CJ * cL.CK) +
                    (lv.aflag ? 1 : 0));
    } else if (kernel == 8) {
      // the last post-sweep of level 0 as the mixed solve runs it with the
      // K-cycle (rho partials with the flexible beta: q read too), register
      // march on the level's FP16 copy
      build_hier32(C);
      ensure_krylov(C, C->k32, false);
      const Level<float> lv = C->h32.L[0];
      const float *qf = C->k32.q;
      MarchShape m;
      if (!lv.h16() || !lv.d || !vec_ok<float>(g) || !march_ok<4>(g, m)) return PSB_EINVAL;
      zero(C, lv.e[0], g.N);
      zero(C, lv.r, g.N);
      zero(C, C->k32.q, g.N);
      DevState h;
      std::memset(&h, 0, sizeof(h));
      HIPCHK(hipMemcpyAsync(C->st32, &h, sizeof(h), hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
      const ProlX<float, 4> pr{};
      const float zs = (float)lv.scale;
      f = [=]() {
        hipLaunchKernelGGL((k_jacobi_march<float, 4, false, true, _Float16, true>),
                           dim3(m.nblocks), dim3(kBlock), 0, stream, g, lv.mat16(), lv.r,
                           lv.e[0], lv.e[1], pr, 0.9f, lv.aflag, C->st32, C->work, zs, m.TI,
                           m.ntiles, qf);
      };
      // four FP16 coefficient arrays, r, e and q in, e out
      *bytes = N * (4 * 2 + 4 * 4 + (lv.aflag ? 1 : 0));
    } else if (kernel == 9) {
      // the FP32 CG x/r step of the mixed solve (x, r, p, q, c in; x, r and
      // the next cycle's first sweep out)
      build_hier32(C);
      ensure_krylov(C, C->k32, false);
      const Level<float> lv = C->h32.L[0];
      const KrylovVecs<float> K = C->k32;
      if (!vec_ok<float>(g)) return PSB_EINVAL;
      zero(C, K.x, g.N);
      zero(C, K.r, g.N);
      zero(C, K.p[0], g.N);
      zero(C, K.q, g.N);
      DevState h;
      std::memset(&h, 0, sizeof(h));
      h.norm = 1;
      HIPCHK(hipMemcpyAsync(C->st32, &h, sizeof(h), hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
      const Metric mt = metric(C);
      f = [=]() {
        hipLaunchKernelGGL((k_cg_xr<float, 4, false>), dim3(nblocks<float>(g)), dim3(kBlock), 0,
                           stream, g, lv.c, mt, K.p[0], K.q, K.x, K.r, lv.e[0], 0.9f, C->st32,
                           C->work);
      };
      *bytes = N * 8 * 4;
    } else if (kernel == 3) {
      ensure32(C);
      ensure_krylov(C, C->k32, false);
      float *x = C->k32.p[0], *y = C->k32.q;
      zero(C, x, g.N);
      const int n = g.N;
      f = [=]() {
        hipLaunchKernelGGL(k_copy, dim3(grid_blocks(n / 4)), dim3(kBlock), 0, stream, x, y, n);
      };
      *bytes = N * (4 + 4);
    } else {
      return PSB_EINVAL;
    }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    f();
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipEventRecord(e0, stream));
    for (int q = 0; q < reps; q++) f();
    HIPCHK(hipEventRecord(e1, stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return PSB_OK;
  });
}

int psb_profile_solve_device(psb_ctx *C, double *x, const double *b, const psb_stop *p,
                             psb_result *R, psb_kernel_stat *stats, double *solve_ms) {
  if (!C || !x || !b || !p || !R || !stats || !solve_ms || !stop_ok(p)) return PSB_EINVAL;
  if (!C->have_matrix) return PSB_ENOTREADY;
  if (p->max_iter < 0 || p->scale == 0.0 || p->stalecount > kStaleMax) return PSB_EINVAL;
  std::memset(R, 0, sizeof(*R));
  std::memset(stats, 0, sizeof(psb_kernel_stat) * PSB_K_COUNT);
  auto clear = [&]() {
    for (auto &t : C->probe.rec) {
      (void)hipEventDestroy(std::get<1>(t));
      (void)hipEventDestroy(std::get<2>(t));
    }
    C->probe.rec.clear();
    C->probe.on = false;
  };
  hipEvent_t t0 = nullptr, t1 = nullptr;
  const int rc = guarded(C, [&]() {
    HIPCHK(hipEventCreate(&t0));
    HIPCHK(hipEventCreate(&t1));
    clear();
    for (double &v : C->probe.bytes) v = 0.0;
    C->probe.on = true;
    HIPCHK(hipEventRecord(t0, C->stream));
    const int r = solve_device(C, x, b, p, R);
    HIPCHK(hipEventRecord(t1, C->stream));
    HIPCHK(hipEventSynchronize(t1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, t0, t1));
    *solve_ms = ms;
    for (auto &t : C->probe.rec) {
      float k = 0.0f;
      HIPCHK(hipEventElapsedTime(&k, std::get<1>(t), std::get<2>(t)));
      psb_kernel_stat &S = stats[std::get<0>(t)];
      S.launches += 1;
      S.ms += k;
    }
    for (int q = 0; q < PSB_K_COUNT; q++) stats[q].bytes = C->probe.bytes[q];
    return r;
  });
  clear();
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  return rc;
}

}  // extern "C"
