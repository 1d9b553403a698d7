// psb_batch.hip -- many independent pressure systems advanced together.
//
// Each system keeps its own context (matrix, geometry, off mask, FP64/FP32
// additive-correction hierarchy, Krylov vectors).  The batch adds device
// tables of per-system pointers and launch lists: every phase of an
// iteration of the mixed-precision AC-PCG with reliable updates (the method
// of psb_solve with PSB_MIXED and inner_adapt = 2) is ONE launch over all
// systems that take part in it, the whole iteration sequence is captured
// once in a hipGraph, and each system stops on its own device-side flag.
// Systems have different depths of hierarchy: at level l, the systems whose
// coarsest level is l apply their dense coarse operator while the others
// smooth and restrict, so one V-cycle of the batch is one descent and one
// ascent over the deepest hierarchy.
//
// The tile of each (level, size class) is picked by a small autotuner that
// times the Jacobi sweep of that class with each candidate: the span form
// with 1, 2, 4 or 8 cell vectors per thread, or (levels of at least 2^15
// cells with 4-wide vectors) the marching form, tiles of 256 vectors of a
// (j, k) plane walking slabs of 4, 8, 16 or 32 i planes (the z-blocking).
// With PSB_TUNE_CACHE naming a file, the choices are read from it when it
// holds every class of the batch (same device) and written back after
// tuning, as JSON.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "psb_bkernels.h"
#include "psb_internal.h"

using namespace psb;
using namespace psb::host;

namespace {

constexpr int kCoarseMinBatch = 4;  // coarsest dims 4..7: dense operators <= 343 cells
constexpr int kVptCand[] = {1, 2, 4, 8};
constexpr int kTiCand[] = {4, 8, 16, 32};  // planes per slab of the marching tile
constexpr int kVptDefault = 2;

struct SegList {
  std::vector<BSeg> segs;  // host copy
  int nblocks = 0;
  BSeg *d_segs = nullptr;
  uint16_t *d_bmap = nullptr;
  bool empty() const { return segs.empty(); }
};

}  // namespace

struct psb_batch {
  int nsys = 0;
  std::vector<psb_ctx *> sys;
  std::vector<int> seen;      // context versions the plan was built for
  hipStream_t stream = nullptr;
  Arena base;                 // states, flags
  Arena arena;                // plan: tables, lists, partials
  DevState *st = nullptr, *hst = nullptr;
  int *flags = nullptr, *hflags = nullptr;
  double *partial = nullptr;
  bool built = false;
  int tune = 0;               // 0: autotune, > 0: fixed vectors per thread
  int V = 1, V2 = 1;          // FP32 / FP64 vector widths of the fine level (all systems)
  std::vector<int> Vl;        // FP32 vector width per level (4 where every system's nz allows)
  int nlev = 0;
  std::vector<int> last;      // coarsest level per system
  std::vector<BFine> fine;
  BFine *d_fine = nullptr;
  std::vector<std::vector<BLev<float>>> lev;  // [level][system]
  std::vector<BLev<float> *> d_lev;
  std::vector<std::vector<int>> vpt;          // [level][system] vectors per thread
  std::vector<std::vector<int>> ti;           // [level][system] planes per slab (0: span,
                                              // -1: as many as fill the segment)
  SegList fineL;                                // all systems, fine level (FP32 vectors)
  SegList fine64L;                              // all systems, fine level (FP64 vectors)
  std::vector<SegList> smoothL;                 // [l]: systems with l < last
  std::vector<std::vector<std::pair<int, SegList>>> restrictL;  // [l]: (CI,CJ,CK) groups
  std::vector<SegList> gemvL;                   // [l]: systems with last == l
  Graph graph;
  double omega = 0, cfac = 0;
  int kmask = 0;                                // K-cycle levels of the tables (bit l: level l)
  bool flex = false;                            // flexible beta in the outer CG (with K-cycles)
  bool kset = false;                            // the tables carry kmask's K-cycle buffers
  std::vector<int> tuned;                       // (level, class, choice) triples:
                                                // choice > 0 vectors per thread (span),
                                                // < 0 minus the planes per slab (march)
};

namespace {

int round8(long long n) { return (int)std::max<long long>(8, (n + 7) / 8 * 8); }

// blocks for `nwork` thread items at `vpt` items per thread
int seg_blocks(long long nwork, int vpt) {
  return round8((nwork + (long long)kBlock * vpt - 1) / ((long long)kBlock * vpt));
}

void upload_list(psb_batch *B, SegList &L) {
  std::vector<uint16_t> bmap;
  int b0 = 0;
  for (size_t q = 0; q < L.segs.size(); q++) {
    L.segs[q].b0 = b0;
    for (int t = 0; t < L.segs[q].nb; t++) bmap.push_back((uint16_t)q);
    b0 += L.segs[q].nb;
  }
  L.nblocks = b0;
  if (L.segs.empty()) return;
  if (L.segs.size() > 65535) throw std::invalid_argument("too many segments in a batch");
  L.d_segs = B->arena.alloc<BSeg>(L.segs.size());
  L.d_bmap = B->arena.alloc<uint16_t>(bmap.size());
  HIPCHK(hipMemcpyAsync(L.d_segs, L.segs.data(), L.segs.size() * sizeof(BSeg),
                        hipMemcpyHostToDevice, B->stream));
  HIPCHK(hipMemcpyAsync(L.d_bmap, bmap.data(), bmap.size() * 2, hipMemcpyHostToDevice,
                        B->stream));
  HIPCHK(hipStreamSynchronize(B->stream));  // bmap is a local buffer
}

// segment of system s on level l with `nwork` thread items
BSeg make_seg(int s, long long nwork, int vpt) { return BSeg{s, 0, seg_blocks(nwork, vpt), 0}; }

long long nvec(const Grid &g, int V) { return (long long)g.N / V; }

int combo_of(const BLev<float> &H) { return (H.sI << 2) | (H.sJ << 1) | H.sK; }

// size class of a level: log2 of its cell count
int size_class(const Grid &g) {
  int c = 0;
  while ((1LL << (c + 1)) <= (long long)g.N) c++;
  return c;
}

constexpr int kBatchMarchMin = 1 << 15;

// whether level l of system s can take the marching tile
bool march_able(const psb_batch *B, int l, int s) {
  const Grid &g = B->lev[l][s].g;
  return B->Vl[l] == 4 && g.N >= kBatchMarchMin && g.nz % 4 == 0;
}

// (slab, tile) items of the marching tile with TI planes per slab
long long march_items(const Grid &g, int TI) {
  const long long ntiles = ((long long)g.ny * (g.nz / 4) + kBlock - 1) / kBlock;
  return ntiles * ((g.nx + TI - 1) / TI);
}

// the segment of system s's streaming kernels on level l: one block per
// marching item when the level marches with a chosen slab depth, else
// vpt vectors per thread
BSeg level_seg(const psb_batch *B, int l, int s, long long nwork) {
  const int TI = B->ti[l][s];
  if (TI > 0) return BSeg{s, 0, round8(march_items(B->lev[l][s].g, TI)), 0};
  return make_seg(s, nwork, B->vpt[l][s]);
}

// launch lists for the current per-(level, system) tiles
void build_lists(psb_batch *B) {
  const int ns = B->nsys;
  B->fineL = SegList();
  B->fine64L = SegList();
  for (int s = 0; s < ns; s++) {
    // the fine list partitions every fine-level kernel that leaves partials
    B->fineL.segs.push_back(level_seg(B, 0, s, nvec(B->fine[s].g, B->V)));
  }
  B->fine64L = B->fineL;  // same partition: partial ranges stay per system
  upload_list(B, B->fineL);
  upload_list(B, B->fine64L);
  B->smoothL.assign(B->nlev, SegList());
  B->gemvL.assign(B->nlev, SegList());
  B->restrictL.assign(B->nlev, {});
  for (int l = 0; l < B->nlev; l++) {
    std::map<int, SegList> groups;
    for (int s = 0; s < ns; s++) {
      if (l < B->last[s]) {
        const BLev<float> &L = B->lev[l][s], &H = B->lev[l + 1][s];
        if (l > 0) B->smoothL[l].segs.push_back(level_seg(B, l, s, nvec(L.g, B->Vl[l])));
        const int KE = B->Vl[l] == 1 ? (H.sK ? 2 : 1) : B->Vl[l];
        const long long nw = (long long)H.g.nx * H.g.ny * (L.g.nz / KE);
        groups[combo_of(H)].segs.push_back(make_seg(s, nw, B->vpt[l][s]));
      } else if (l == B->last[s]) {
        const int rows = B->lev[l][s].m;
        B->gemvL[l].segs.push_back(BSeg{s, 0, round8((rows + kWaves - 1) / kWaves), 0});
      }
    }
    if (l == 0) B->smoothL[0] = B->fineL;  // level 0: the fine partition (RHO partials)
    else upload_list(B, B->smoothL[l]);
    upload_list(B, B->gemvL[l]);
    for (auto &kv : groups) {
      upload_list(B, kv.second);
      B->restrictL[l].push_back(kv);
    }
  }
  int most = B->fineL.nblocks;
  for (const SegList &L : B->smoothL) most = std::max(most, L.nblocks);
  B->partial = B->arena.alloc<double>((size_t)most * kMaxPartials);
}

// the marching fields of the level tables (the Jacobi sweeps, and on level
// 0 the CG p/q step, march where TI > 0): the chosen slab depth, or (ti ==
// -1, a fixed work per thread) slabs sized so that the system's segment
// holds about one tile-slab per block; span where ti == 0
void march_tables(psb_batch *B) {
  for (int l = 0; l < B->nlev; l++) {
    const SegList &L = l == 0 ? B->fineL : B->smoothL[l];
    for (const BSeg &sg : L.segs) {
      BLev<float> &D = B->lev[l][sg.sys];
      D.TI = D.ntiles = 0;
      const int want = B->ti[l][sg.sys];
      if (want == 0 || !march_able(B, l, sg.sys)) continue;
      const long long nvp = (long long)D.g.ny * (D.g.nz / 4);
      const int ntiles = (int)((nvp + kBlock - 1) / kBlock);
      int TI = want;
      if (TI < 0) TI = std::max(1, (int)(((long long)D.g.nx * ntiles + sg.nb - 1) / sg.nb));
      D.TI = TI;
      D.ntiles = ntiles;
    }
    HIPCHK(hipMemcpyAsync(B->d_lev[l], B->lev[l].data(), sizeof(BLev<float>) * B->nsys,
                          hipMemcpyHostToDevice, B->stream));
  }
  HIPCHK(hipStreamSynchronize(B->stream));
}

// ---------------------------------------------------------------------------
// launchers over a list
// ---------------------------------------------------------------------------
template <bool PROL, bool RHO>
void l_jacobi(psb_batch *B, int l, int cur, int nxt, int outH) {
  const SegList &L = B->smoothL[l];
  if (L.empty()) return;
  const Work w{B->partial};
  const BLev<float> *lh = l + 1 < B->nlev ? B->d_lev[l + 1] : nullptr;
  const float cfac = (float)B->cfac, omega = (float)B->omega;
  const bool flex = RHO && B->flex;
  if (B->Vl[l] == 4)
    hipLaunchKernelGGL((kb_jacobi<float, 4, PROL, RHO>), dim3(L.nblocks), dim3(kBlock), 0,
                       B->stream, L.d_segs, L.d_bmap, B->d_fine, B->d_lev[l], lh, cur, nxt, outH,
                       cfac, omega, w, flex);
  else
    hipLaunchKernelGGL((kb_jacobi<float, 1, PROL, RHO>), dim3(L.nblocks), dim3(kBlock), 0,
                       B->stream, L.d_segs, L.d_bmap, B->d_fine, B->d_lev[l], lh, cur, nxt, outH,
                       cfac, omega, w, flex);
}

template <int V, int KV, int CI, int CJ, int CK>
void l_restrict_v(psb_batch *B, const SegList &L, int l, int cur) {
  hipLaunchKernelGGL((kb_restrict<float, V, KV, CI, CJ, CK>), dim3(L.nblocks), dim3(kBlock), 0,
                     B->stream, L.d_segs, L.d_bmap, B->d_fine, B->d_lev[l], B->d_lev[l + 1], cur,
                     (float)B->omega);
}

template <int CI, int CJ, int CK>
void l_restrict_c(psb_batch *B, const SegList &L, int l, int cur) {
  if (B->Vl[l] == 4) {
    l_restrict_v<4, 1, CI, CJ, CK>(B, L, l, cur);
  } else {
    if constexpr (CK == 2) l_restrict_v<1, 2, CI, CJ, CK>(B, L, l, cur);
    else l_restrict_v<1, 1, CI, CJ, CK>(B, L, l, cur);
  }
}

void l_restrict(psb_batch *B, int l, int cur) {
  for (auto &g : B->restrictL[l]) {
    const SegList &L = g.second;
    switch (g.first) {
      case 7: l_restrict_c<2, 2, 2>(B, L, l, cur); break;
      case 6: l_restrict_c<2, 2, 1>(B, L, l, cur); break;
      case 5: l_restrict_c<2, 1, 2>(B, L, l, cur); break;
      case 4: l_restrict_c<2, 1, 1>(B, L, l, cur); break;
      case 3: l_restrict_c<1, 2, 2>(B, L, l, cur); break;
      case 2: l_restrict_c<1, 2, 1>(B, L, l, cur); break;
      case 1: l_restrict_c<1, 1, 2>(B, L, l, cur); break;
      default: throw std::invalid_argument("level without coarsening");
    }
  }
}

template <int NR>
void l_finalize(psb_batch *B, int op) {
  hipLaunchKernelGGL((kb_finalize<NR>), dim3(B->nsys), dim3(kBlock), 0, B->stream, op,
                     B->fineL.d_segs, B->d_fine, B->partial, B->flags);
}

void l_ru_resid(psb_batch *B, bool gated, int op) {
  const SegList &L = B->fine64L;
  const Work w{B->partial};
  if (B->V2 == 2)
    hipLaunchKernelGGL((kb_ru_resid<2>), dim3(L.nblocks), dim3(kBlock), 0, B->stream, L.d_segs,
                       L.d_bmap, B->d_fine, B->d_lev[0], (float)B->omega, gated, w);
  else
    hipLaunchKernelGGL((kb_ru_resid<1>), dim3(L.nblocks), dim3(kBlock), 0, B->stream, L.d_segs,
                       L.d_bmap, B->d_fine, B->d_lev[0], (float)B->omega, gated, w);
  l_finalize<4>(B, op);
}

void l_ru_xupdate(psb_batch *B, bool gated) {
  const SegList &L = B->fine64L;
  if (B->V2 == 2)
    hipLaunchKernelGGL((kb_ru_xupdate<2>), dim3(L.nblocks), dim3(kBlock), 0, B->stream,
                       L.d_segs, L.d_bmap, B->d_fine, gated);
  else
    hipLaunchKernelGGL((kb_ru_xupdate<1>), dim3(L.nblocks), dim3(kBlock), 0, B->stream,
                       L.d_segs, L.d_bmap, B->d_fine, gated);
}

// K-cycle steps of level l over the systems that smooth there
template <typename F>
void l_kc(psb_batch *B, int l, F launch) {
  const SegList &L = B->smoothL[l];
  if (L.empty()) return;
  if (B->Vl[l] == 4) launch(std::integral_constant<int, 4>(), L);
  else launch(std::integral_constant<int, 1>(), L);
}

void b_cycle(psb_batch *B, int l, int nu1, int nu2);

// level l's correction by two flexible-CG steps preconditioned by the cycle
// on l (as kstep in the single-system solver), in kz of the systems that
// smooth on l; the systems whose coarsest level is l keep their dense solve
void b_kstep(psb_batch *B, int l, int nu1, int nu2) {
  const int out = (nu1 - 1 + nu2) % 3;
  const float omega = (float)B->omega;
  b_cycle(B, l, nu1, nu2);
  l_kc(B, l, [&](auto vt, const SegList &L) {
    constexpr int V = decltype(vt)::value;
    hipLaunchKernelGGL((kb_kc_av<float, V>), dim3(L.nblocks), dim3(kBlock), 0, B->stream,
                       L.d_segs, L.d_bmap, B->d_fine, B->d_lev[l], out, Work{B->partial});
    hipLaunchKernelGGL((kb_kc_r<float, V>), dim3(L.nblocks), dim3(kBlock), 0, B->stream,
                       L.d_segs, L.d_bmap, B->d_fine, B->d_lev[l], omega, B->partial);
  });
  b_cycle(B, l, nu1, nu2);
  l_kc(B, l, [&](auto vt, const SegList &L) {
    constexpr int V = decltype(vt)::value;
    hipLaunchKernelGGL((kb_kc_dw<float, V>), dim3(L.nblocks), dim3(kBlock), 0, B->stream,
                       L.d_segs, L.d_bmap, B->d_fine, B->d_lev[l], out, Work{B->partial});
    hipLaunchKernelGGL((kb_kc_comb<float, V>), dim3(L.nblocks), dim3(kBlock), 0, B->stream,
                       L.d_segs, L.d_bmap, B->d_fine, B->d_lev[l], out, B->partial);
  });
}

// the cycle on level l of every system that reaches it (its correction in
// e[out]; the first sweep already in e[0]): the dense solve where l is the
// coarsest level, else pre-sweeps, restriction, the coarse correction (the
// K-cycle's two steps on K levels) and the post-sweeps with its prolongation
void b_cycle(psb_batch *B, int l, int nu1, int nu2) {
  const int out = (nu1 - 1 + nu2) % 3;
  const SegList &G = B->gemvL[l];
  if (!G.empty())
    hipLaunchKernelGGL((kb_gemv<float>), dim3(G.nblocks), dim3(kBlock), 0, B->stream, G.d_segs,
                       G.d_bmap, B->d_fine, B->d_lev[l], out);
  if (B->smoothL[l].empty()) return;
  int cur = 0;
  for (int s = 1; s < nu1; s++) {
    const int nxt = (cur + 1) % 3;
    l_jacobi<false, false>(B, l, cur, nxt, 0);
    cur = nxt;
  }
  l_restrict(B, l, cur);
  if ((B->kmask >> (l + 1)) & 1) b_kstep(B, l + 1, nu1, nu2);
  else b_cycle(B, l + 1, nu1, nu2);
  for (int s = 0; s < nu2; s++) {
    const int nxt = (cur + 1) % 3;
    const bool last = (s == nu2 - 1) && l == 0;
    if (s == 0 && last) l_jacobi<true, true>(B, l, cur, nxt, out);
    else if (s == 0) l_jacobi<true, false>(B, l, cur, nxt, out);
    else if (last) l_jacobi<false, true>(B, l, cur, nxt, out);
    else l_jacobi<false, false>(B, l, cur, nxt, out);
    cur = nxt;
  }
}

// one preconditioner application of every system (z in level 0's e[out])
// and the rho step
void b_vcycle(psb_batch *B, int nu1, int nu2) {
  b_cycle(B, 0, nu1, nu2);
  l_finalize<kRhoParts>(B, B->flex ? FIN_RHO_FLEX : FIN_RHO);
}

void b_iteration(psb_batch *B, int par, int nu1, int nu2) {
  const int out = (nu1 - 1 + nu2) % 3;
  b_vcycle(B, nu1, nu2);
  const SegList &L = B->fineL;
  const Work w{B->partial};
  if (B->V == 4) {
    hipLaunchKernelGGL((kb_cg_pq<4>), dim3(L.nblocks), dim3(kBlock), 0, B->stream, L.d_segs,
                       L.d_bmap, B->d_fine, B->d_lev[0], out, par, w);
  } else {
    hipLaunchKernelGGL((kb_cg_pq<1>), dim3(L.nblocks), dim3(kBlock), 0, B->stream, L.d_segs,
                       L.d_bmap, B->d_fine, B->d_lev[0], out, par, w);
  }
  l_finalize<1>(B, FIN_PQ);
  if (B->V == 4) {
    hipLaunchKernelGGL((kb_cg_xr<4>), dim3(L.nblocks), dim3(kBlock), 0, B->stream, L.d_segs,
                       L.d_bmap, B->d_fine, B->d_lev[0], par, (float)B->omega, w);
  } else {
    hipLaunchKernelGGL((kb_cg_xr<1>), dim3(L.nblocks), dim3(kBlock), 0, B->stream, L.d_segs,
                       L.d_bmap, B->d_fine, B->d_lev[0], par, (float)B->omega, w);
  }
  l_finalize<3>(B, FIN_XR_RU);
  l_ru_xupdate(B, true);
  l_ru_resid(B, true, FIN_RU);
}

// ---------------------------------------------------------------------------
// plan: tables, lists, autotuning
// ---------------------------------------------------------------------------
bool stale(psb_batch *B) {
  if (!B->built) return true;
  for (int s = 0; s < B->nsys; s++)
    if (B->sys[s]->version != B->seen[s]) return true;
  return false;
}

// time the level-l Jacobi sweep of the systems in `who` with the span tile
// (vpt vectors per thread, TI = 0) or the marching tile (TI planes per slab)
double time_class(psb_batch *B, int l, const std::vector<int> &who, int vpt, int TI) {
  SegList L;
  for (int s : who) {
    BLev<float> &D = B->lev[l][s];
    if (TI > 0) {
      D.TI = TI;
      D.ntiles = (int)(((long long)D.g.ny * (D.g.nz / 4) + kBlock - 1) / kBlock);
      L.segs.push_back(BSeg{s, 0, round8(march_items(D.g, TI)), 0});
    } else {
      D.TI = D.ntiles = 0;
      L.segs.push_back(make_seg(s, nvec(D.g, B->Vl[l]), vpt));
    }
  }
  HIPCHK(hipMemcpyAsync(B->d_lev[l], B->lev[l].data(), sizeof(BLev<float>) * B->nsys,
                        hipMemcpyHostToDevice, B->stream));
  upload_list(B, L);
  const BLev<float> *lv = B->d_lev[l];
  const float omega = 0.9f;
  auto launch = [&]() {
    if (B->Vl[l] == 4)
      hipLaunchKernelGGL((kb_jacobi<float, 4, false, false>), dim3(L.nblocks), dim3(kBlock), 0,
                         B->stream, L.d_segs, L.d_bmap, B->d_fine, lv, lv, 0, 1, 0, 0.0f, omega,
                         Work{B->partial}, false);
    else
      hipLaunchKernelGGL((kb_jacobi<float, 1, false, false>), dim3(L.nblocks), dim3(kBlock), 0,
                         B->stream, L.d_segs, L.d_bmap, B->d_fine, lv, lv, 0, 1, 0, 0.0f, omega,
                         Work{B->partial}, false);
  };
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  launch();
  HIPCHK(hipEventRecord(e0, B->stream));
  const int reps = 5;
  for (int q = 0; q < reps; q++) launch();
  HIPCHK(hipEventRecord(e1, B->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}

// ---------------------------------------------------------------------------
// tuning cache (PSB_TUNE_CACHE): {"device": "<arch>:<CUs>", "entries": [
// {"level": l, "class": c, "choice": v, "ms": t}, ...]} -- one entry per
// (level, size class); choice as in psb_batch::tuned
// ---------------------------------------------------------------------------
std::string device_tag() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipDeviceProp_t pr;
  HIPCHK(hipGetDeviceProperties(&pr, dev));
  return std::string(pr.gcnArchName) + ":" + std::to_string(pr.multiProcessorCount);
}

struct TuneEntry {
  int level, cls, choice;
  double ms;
};

// the entries of the cache file for this device (empty if absent or for
// another device)
std::vector<TuneEntry> read_cache(const char *path, const std::string &tag) {
  std::vector<TuneEntry> out;
  std::ifstream f(path);
  if (!f) return out;
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string t = ss.str();
  const std::string dkey = "\"device\": \"" + tag + "\"";
  if (t.find(dkey) == std::string::npos) return out;
  size_t pos = 0;
  while ((pos = t.find("{\"level\":", pos)) != std::string::npos) {
    TuneEntry e{};
    if (std::sscanf(t.c_str() + pos, "{\"level\": %d, \"class\": %d, \"choice\": %d, \"ms\": %lf",
                    &e.level, &e.cls, &e.choice, &e.ms) == 4)
      out.push_back(e);
    pos++;
  }
  return out;
}

void write_cache(const char *path, const std::string &tag, const std::vector<TuneEntry> &es) {
  std::ofstream f(path);
  if (!f) return;  // a cache that cannot be written only costs the next tuning
  f << "{\"device\": \"" << tag << "\", \"entries\": [";
  for (size_t q = 0; q < es.size(); q++) {
    char buf[160];
    std::snprintf(buf, sizeof(buf), "%s\n  {\"level\": %d, \"class\": %d, \"choice\": %d, \"ms\": %.6g}",
                  q ? "," : "", es[q].level, es[q].cls, es[q].choice, es[q].ms);
    f << buf;
  }
  f << "\n]}\n";
}

void autotune(psb_batch *B) {
  B->tuned.clear();
  // the tuning sweeps need the states to read "not done"
  std::vector<DevState> h(B->nsys);
  std::memset(h.data(), 0, sizeof(DevState) * B->nsys);
  HIPCHK(hipMemcpyAsync(B->st, h.data(), sizeof(DevState) * B->nsys, hipMemcpyHostToDevice,
                        B->stream));
  const char *cache = std::getenv("PSB_TUNE_CACHE");
  const std::string tag = cache ? device_tag() : std::string();
  std::vector<TuneEntry> known = cache ? read_cache(cache, tag) : std::vector<TuneEntry>();
  std::map<std::pair<int, int>, TuneEntry> table;
  for (const TuneEntry &e : known) table[{e.level, e.cls}] = e;
  bool fresh = false;
  for (int l = 0; l < B->nlev; l++) {
    std::map<int, std::vector<int>> classes;
    for (int s = 0; s < B->nsys; s++)
      if (l < B->last[s]) classes[size_class(B->lev[l][s].g)].push_back(s);
    for (auto &kv : classes) {
      auto hit = table.find({l, kv.first});
      int best = kVptDefault;
      if (hit != table.end()) {
        best = hit->second.choice;
      } else {
        double tbest = 1e300;
        for (int v : kVptCand) {
          const double t = time_class(B, l, kv.second, v, 0);
          if (t < tbest) {
            tbest = t;
            best = v;
          }
        }
        bool marchable = true;
        for (int s : kv.second) marchable = marchable && march_able(B, l, s);
        if (marchable) {
          for (int TI : kTiCand) {
            const double t = time_class(B, l, kv.second, 0, TI);
            if (t < tbest) {
              tbest = t;
              best = -TI;
            }
          }
        }
        table[{l, kv.first}] = TuneEntry{l, kv.first, best, tbest};
        fresh = true;
      }
      bool marchable = best < 0;
      for (int s : kv.second) marchable = marchable && march_able(B, l, s);
      for (int s : kv.second) {
        B->vpt[l][s] = best > 0 ? best : kVptDefault;
        B->ti[l][s] = marchable ? -best : 0;
      }
      B->tuned.push_back(l);
      B->tuned.push_back(kv.first);
      B->tuned.push_back(marchable ? best : B->vpt[l][kv.second[0]]);
    }
  }
  HIPCHK(hipStreamSynchronize(B->stream));
  if (cache && fresh) {
    std::vector<TuneEntry> all;
    for (auto &kv : table) all.push_back(kv.second);
    write_cache(cache, tag, all);
  }
}

void build(psb_batch *B) {
  HIPCHK(hipStreamSynchronize(B->stream));
  destroy_graph(B->graph);
  B->arena.release();
  B->built = false;
  const int ns = B->nsys;
  // per-system set-up on each context's own stream
  for (int s = 0; s < ns; s++) {
    psb_ctx *C = B->sys[s];
    if (!C->have_matrix) throw std::invalid_argument("psb_set_matrix not called for a system");
    if (!C->sym) throw std::invalid_argument("batched solves need symmetric matrices");
    build_hier32(C);
    HIPCHK(hipStreamSynchronize(C->stream));
  }
  B->nlev = 0;
  B->last.assign(ns, 0);
  bool v2 = true;
  for (int s = 0; s < ns; s++) {
    const Hier<float> &H = B->sys[s]->h32;
    B->last[s] = (int)H.L.size() - 1;
    B->nlev = std::max(B->nlev, (int)H.L.size());
    v2 = v2 && B->sys[s]->g.nz % 2 == 0;
  }
  B->Vl.assign(B->nlev, 4);
  for (int s = 0; s < ns; s++) {
    const Hier<float> &H = B->sys[s]->h32;
    for (int l = 0; l < (int)H.L.size(); l++)
      if (H.L[l].g.nz % 4 != 0) B->Vl[l] = 1;
  }
  B->V = B->Vl[0];
  B->V2 = v2 ? 2 : 1;
  // tables
  B->lev.assign(B->nlev, std::vector<BLev<float>>(ns));
  B->fine.assign(ns, BFine{});
  for (int s = 0; s < ns; s++) {
    psb_ctx *C = B->sys[s];
    const Hier<float> &H = C->h32;
    for (int l = 0; l < (int)H.L.size(); l++) {
      const Level<float> &L = H.L[l];
      BLev<float> &D = B->lev[l][s];
      D.g = L.g;
      D.c = L.c; D.w = L.w; D.s = L.s; D.b = L.b;
      D.d = L.d;
      if (!L.d) throw std::invalid_argument("batched solves need the flux-form levels");
      D.r = L.r;
      for (int q = 0; q < 3; q++) D.e[q] = L.e[q];
      D.af = L.aflag;
      D.sI = L.CI == 2 ? 1 : 0;
      D.sJ = L.CJ == 2 ? 1 : 0;
      D.sK = L.CK == 2 ? 1 : 0;
      D.m = l == B->last[s] ? H.m : 0;
      D.G = l == B->last[s] ? H.G : nullptr;
    }
    BFine &F = B->fine[s];
    F.g = C->g;
    F.A = mat64(C);
    F.met = metric(C);
    F.d = C->k32.x;
    F.r = C->k32.r;
    F.p[0] = C->k32.p[0];
    F.p[1] = C->k32.p[1];
    F.q = C->k32.q;
    F.st = B->st + s;
  }
  B->d_lev.assign(B->nlev, nullptr);
  for (int l = 0; l < B->nlev; l++) {
    B->d_lev[l] = B->arena.alloc<BLev<float>>(ns);
    HIPCHK(hipMemcpyAsync(B->d_lev[l], B->lev[l].data(), sizeof(BLev<float>) * ns,
                          hipMemcpyHostToDevice, B->stream));
  }
  B->d_fine = B->arena.alloc<BFine>(ns);
  B->vpt.assign(B->nlev, std::vector<int>(ns, B->tune > 0 ? B->tune : kVptDefault));
  // a fixed work per thread: the large levels march with slabs that fill
  // their segments; the autotuner chooses per (level, class) otherwise
  B->ti.assign(B->nlev, std::vector<int>(ns, -1));
  // the fine tables point at the callers' x and b only at solve time
  HIPCHK(hipMemcpyAsync(B->d_fine, B->fine.data(), sizeof(BFine) * ns, hipMemcpyHostToDevice,
                        B->stream));
  B->partial = B->arena.alloc<double>((size_t)ns * 8 * kMaxPartials);
  HIPCHK(hipStreamSynchronize(B->stream));
  if (B->tune == 0) autotune(B);
  build_lists(B);
  march_tables(B);
  HIPCHK(hipStreamSynchronize(B->stream));
  for (int s = 0; s < ns; s++) B->seen[s] = B->sys[s]->version;
  B->built = true;
  B->kset = false;  // new level tables: the K-cycle fields are set per solve
}

// K-cycle levels of a batch: by level index (the hierarchies differ in
// depth), on every system that smooths on that level; -1 chooses levels 1
// and 2
constexpr int kBatchKcycleAuto = 0b110;

// the level tables' K-cycle fields for `mask`: buffers from each system's
// hierarchy arena (kept with the hierarchy), flags on the levels a system
// smooths on
void set_kcycle(psb_batch *B, int mask) {
  if (B->kset && B->kmask == mask) return;
  for (int l = 0; l < B->nlev; l++) {
    for (int s = 0; s < B->nsys; s++) {
      BLev<float> &D = B->lev[l][s];
      D.kcyc = l >= 1 && ((mask >> l) & 1) && l < B->last[s];
      if (!D.kcyc) continue;
      Level<float> &lv = B->sys[s]->h32.L[l];
      if (!lv.kc) {
        psb_ctx *C = B->sys[s];
        lv.kc = C->harena.alloc<float>(lv.g.N);
        lv.kv = C->harena.alloc<float>(lv.g.N);
        lv.kz = C->harena.alloc<float>(lv.g.N);
        lv.kst = C->harena.alloc<double>(4);
      }
      D.kc = lv.kc;
      D.kv = lv.kv;
      D.kz = lv.kz;
      D.kst = lv.kst;
    }
    HIPCHK(hipMemcpyAsync(B->d_lev[l], B->lev[l].data(), sizeof(BLev<float>) * B->nsys,
                          hipMemcpyHostToDevice, B->stream));
  }
  HIPCHK(hipStreamSynchronize(B->stream));
  B->kmask = mask;
  B->kset = true;
}

int batch_solve(psb_batch *B, double *const *x, const double *const *b, const psb_stop *p,
                psb_result *R) {
  if (p->precond != PSB_PREC_AC || p->method != PSB_CG || p->precision != PSB_MIXED ||
      p->inner_adapt != 2 || p->ac_pre < 1 || p->ac_post < 1)
    return PSB_EINVAL;
  if (p->ac_kcycle < -1 || (p->ac_kcycle > 0 && (p->ac_kcycle & 1))) return PSB_EINVAL;
  if (stale(B)) build(B);
  if (p->ac_flexible < -1 || p->ac_flexible > 1) return PSB_EINVAL;
  const int kmask = p->ac_kcycle == -1 ? kBatchKcycleAuto : p->ac_kcycle;
  // the outer CG's beta: flexible when asked, or by default when the
  // preconditioner varies (K-cycle)
  const bool flex = p->ac_flexible == 1 || (p->ac_flexible == -1 && kmask != 0);
  if (!B->kset || kmask != B->kmask || flex != B->flex) destroy_graph(B->graph);
  set_kcycle(B, kmask);
  B->flex = flex;
  const int ns = B->nsys;
  // point the fine tables at this call's arrays
  for (int s = 0; s < ns; s++) {
    if (!x[s] || !b[s]) return PSB_EINVAL;
    B->fine[s].x = x[s];
    B->fine[s].b = b[s];
  }
  HIPCHK(hipMemcpyAsync(B->d_fine, B->fine.data(), sizeof(BFine) * ns, hipMemcpyHostToDevice,
                        B->stream));
  // states
  for (int s = 0; s < ns; s++) {
    DevState &h = B->hst[s];
    std::memset(&h, 0, sizeof(h));
    h.res_rat = p->res_rat;
    h.res_tol = p->res_tol;
    h.scale = p->scale;
    h.ntot = (double)B->sys[s]->g.N;
    h.mini = p->min_iter;
    h.maxit = p->max_iter;
    h.stalecount = -1;
    h.norm = p->norm;
    h.delta = p->inner_rtol;
  }
  HIPCHK(hipMemcpyAsync(B->st, B->hst, sizeof(DevState) * ns, hipMemcpyHostToDevice,
                        B->stream));
  const int nu1 = p->ac_pre, nu2 = p->ac_post;
  const int chunk = std::max(2, p->graph_chunk + (p->graph_chunk & 1));
  const long long key = (long long)chunk + ((long long)nu1 << 16) + ((long long)nu2 << 24);
  if (B->graph.key != key || B->omega != p->ac_omega || B->cfac != p->ac_cfac)
    destroy_graph(B->graph);
  B->omega = p->ac_omega;
  B->cfac = p->ac_cfac;
  {
    const SegList &L = B->fineL;
    if (B->V == 4)
      hipLaunchKernelGGL((kb_zero<4>), dim3(L.nblocks), dim3(kBlock), 0, B->stream, L.d_segs,
                         L.d_bmap, B->d_fine);
    else
      hipLaunchKernelGGL((kb_zero<1>), dim3(L.nblocks), dim3(kBlock), 0, B->stream, L.d_segs,
                         L.d_bmap, B->d_fine);
  }
  l_ru_resid(B, false, FIN_RU_START);
  HIPCHK(hipGetLastError());
  auto all_done = [&]() {
    HIPCHK(hipMemcpyAsync(B->hflags, B->flags, sizeof(int) * ns, hipMemcpyDeviceToHost,
                          B->stream));
    HIPCHK(hipStreamSynchronize(B->stream));
    for (int s = 0; s < ns; s++)
      if (!B->hflags[s]) return false;
    return true;
  };
  long long launches = 0;
  while (!all_done()) {
    if (!B->graph.exec) {
      HIPCHK(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
      for (int it = 0; it < chunk; it++) b_iteration(B, it & 1, nu1, nu2);
      hipError_t e = hipStreamEndCapture(B->stream, &B->graph.graph);
      if (e != hipSuccess) throw HipError{e};
      HIPCHK(hipGraphInstantiate(&B->graph.exec, B->graph.graph, nullptr, nullptr, 0));
      B->graph.key = key;
    }
    HIPCHK(hipGraphLaunch(B->graph.exec, B->stream));
    if (++launches > (long long)p->max_iter + 2) break;
  }
  // fold in the last corrections; the final true residuals give ||b - A x||
  // (||b|| came with the first one)
  l_ru_xupdate(B, false);
  l_ru_resid(B, false, FIN_RU_END);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(B->hst, B->st, sizeof(DevState) * ns, hipMemcpyDeviceToHost,
                        B->stream));
  HIPCHK(hipStreamSynchronize(B->stream));
  int status = PSB_OK;
  for (int s = 0; s < ns; s++) {
    const DevState &h = B->hst[s];
    finish_result(&R[s], h, h.outer);
    R[s].res = h.res_true;
    R[s].rel_l2 = h.bsq > 0.0 ? h.rl2_true / std::sqrt(h.bsq) : 0.0;
    if (status == PSB_OK && R[s].status != PSB_OK) status = R[s].status;
  }
  return status;
}

template <typename F>
int bguard(psb_batch *B, F f) {
  try {
    return f();
  } catch (const HipError &e) {
    return e.e == hipErrorOutOfMemory ? PSB_ENOMEM : PSB_EHIP;
  } catch (const std::bad_alloc &) {
    return PSB_ENOMEM;
  } catch (const std::invalid_argument &) {
    return PSB_EINVAL;
  }
}

}  // namespace

extern "C" {

int psb_batch_create(psb_batch **out, int nsys, const psb_layout *layouts) {
  if (!out || nsys < 1 || !layouts) return PSB_EINVAL;
  *out = nullptr;
  psb_batch *B = new (std::nothrow) psb_batch();
  if (!B) return PSB_ENOMEM;
  B->nsys = nsys;
  B->sys.assign(nsys, nullptr);
  B->seen.assign(nsys, -1);
  int rc = bguard(B, [&]() {
    for (int s = 0; s < nsys; s++) {
      int r = psb_create(&B->sys[s], &layouts[s]);
      if (r != PSB_OK) return r;
      B->sys[s]->coarse_min = kCoarseMinBatch;
    }
    HIPCHK(hipStreamCreateWithFlags(&B->stream, hipStreamDefault));  // orders after the null stream
    B->st = B->base.alloc<DevState>(nsys);
    B->flags = B->base.alloc<int>(nsys);
    HIPCHK(hipHostMalloc((void **)&B->hst, sizeof(DevState) * nsys, hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&B->hflags, sizeof(int) * nsys, hipHostMallocDefault));
    return (int)PSB_OK;
  });
  if (rc != PSB_OK) {
    psb_batch_destroy(B);
    return rc;
  }
  *out = B;
  return PSB_OK;
}

int psb_batch_destroy(psb_batch *B) {
  if (!B) return PSB_OK;
  if (B->stream) (void)hipStreamSynchronize(B->stream);
  destroy_graph(B->graph);
  B->arena.free_all();
  B->base.free_all();
  if (B->hst) (void)hipHostFree(B->hst);
  if (B->hflags) (void)hipHostFree(B->hflags);
  if (B->stream) (void)hipStreamDestroy(B->stream);
  for (psb_ctx *C : B->sys) psb_destroy(C);
  delete B;
  return PSB_OK;
}

psb_ctx *psb_batch_system(psb_batch *B, int sys) {
  if (!B || sys < 0 || sys >= B->nsys) return nullptr;
  return B->sys[sys];
}

int psb_batch_set_tuning(psb_batch *B, int vpt) {
  if (!B || vpt < 0 || vpt > 64) return PSB_EINVAL;
  B->tune = vpt;
  B->built = false;
  return PSB_OK;
}

int psb_batch_plan(psb_batch *B, int *triples, int cap) {
  if (!B) return PSB_EINVAL;
  const int n = (int)B->tuned.size() / 3;
  for (int q = 0; q < std::min(n, cap) * 3; q++) triples[q] = B->tuned[q];
  return n;
}

int psb_batch_solve_device(psb_batch *B, double *const *x, const double *const *b,
                           const psb_stop *p, psb_result *R) {
  if (!B || !x || !b || !p || !R || !stop_ok(p)) return PSB_EINVAL;
  return bguard(B, [&]() { return batch_solve(B, x, b, p, R); });
}

int psb_batch_solve(psb_batch *B, double *const *x, const double *const *b, const psb_stop *p,
                    psb_result *R) {
  if (!B || !x || !b || !p || !R || !stop_ok(p)) return PSB_EINVAL;
  return bguard(B, [&]() {
    const int ns = B->nsys;
    std::vector<double *> dx(ns);
    std::vector<const double *> db(ns);
    std::vector<double> buf;
    for (int s = 0; s < ns; s++) {
      psb_ctx *C = B->sys[s];
      if (!x[s] || !b[s]) return (int)PSB_EINVAL;
      const size_t N = C->g.N;
      gather_host(C->lay, x[s], buf);
      HIPCHK(hipMemcpy(C->dx, buf.data(), N * 8, hipMemcpyHostToDevice));
      gather_host(C->lay, b[s], buf);
      HIPCHK(hipMemcpy(C->db, buf.data(), N * 8, hipMemcpyHostToDevice));
      dx[s] = C->dx;
      db[s] = C->db;
    }
    int rc = batch_solve(B, dx.data(), db.data(), p, R);
    for (int s = 0; s < ns; s++) {
      psb_ctx *C = B->sys[s];
      buf.resize(C->g.N);
      HIPCHK(hipMemcpy(buf.data(), C->dx, buf.size() * 8, hipMemcpyDeviceToHost));
      scatter_host(C->lay, buf, x[s]);
    }
    return rc;
  });
}

}  // extern "C"
