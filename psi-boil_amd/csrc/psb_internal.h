// psb_internal.h -- host-side types shared by the single-system solver
// (psb_api.hip) and the batched solver (psb_batch.hip): device arenas, the
// additive-correction hierarchy, Krylov work vectors, captured graphs and
// the context behind psb_ctx.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <tuple>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/psb.h"
#include "psb_kernels.h"

namespace psb {
namespace host {

struct HipError {
  hipError_t e;
};

#define HIPCHK(x)                               \
  do {                                          \
    hipError_t e_ = (x);                        \
    if (e_ != hipSuccess) throw HipError{e_};   \
  } while (0)

// blocks for a streaming kernel over `nwork` items: one pass of 256 threads
// per item run, at most kMaxBlocks (8 per CU on the 256 CUs), a multiple of
// the 8 XCDs so the XCD remap applies
inline int grid_blocks(long long nwork) {
  long long nb = (nwork + kBlock - 1) / kBlock;
  if (nb > kMaxBlocks) nb = kMaxBlocks;
  if (nb >= 8) nb = (nb / 8) * 8;
  if (nb < 1) nb = 1;
  return (int)nb;
}

inline FastDiv make_div(uint32_t d) {
  uint32_t s = 0;
  while ((1ull << s) < d) s++;
  FastDiv f;
  f.d = d;
  f.m = (uint32_t)(((1ull << (31 + s)) + d - 1) / d);
  f.sh = 31 + s;
  return f;
}

inline Grid make_grid(int nx, int ny, int nz) {
  Grid g;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.sI = ny * nz;
  g.N = nx * ny * nz;
  g.dI = make_div((uint32_t)g.sI);
  g.dZ = make_div((uint32_t)nz);
  return g;
}

// Device allocations of a context.  release() keeps the blocks for reuse:
// a matrix that changes every time step (two-phase flow) is re-assembled and
// its hierarchy rebuilt with the same block sizes, which then cost no
// hipMalloc / hipFree (each of those synchronizes the device and maps or
// unmaps pages).  Blocks left unused for a whole cycle between two releases
// are freed at the second; free_all() returns everything.
struct Arena {
  std::vector<std::pair<void *, size_t>> ptrs;  // live blocks
  std::multimap<size_t, void *> spare;          // released, kept for reuse
  template <typename T>
  T *alloc(size_t n) {
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    bytes = (bytes + 255) & ~size_t(255);
    void *p = nullptr;
    auto it = spare.find(bytes);
    if (it != spare.end()) {
      p = it->second;
      spare.erase(it);
    } else {
      hipError_t e = hipMalloc(&p, bytes);
      if (e != hipSuccess && !spare.empty()) {
        (void)hipGetLastError();
        purge();
        e = hipMalloc(&p, bytes);
      }
      if (e != hipSuccess) throw HipError{e};
    }
    ptrs.emplace_back(p, bytes);
    return static_cast<T *>(p);
  }
  // every caller has synchronized its stream; the device-wide wait covers
  // work of other streams on the blocks (hipFree used to imply it)
  void release() {
    if (ptrs.empty()) return;
    (void)hipDeviceSynchronize();
    purge();
    for (auto &q : ptrs) spare.emplace(q.second, q.first);
    ptrs.clear();
  }
  void purge() {
    for (auto &q : spare) (void)hipFree(q.second);
    spare.clear();
  }
  void free_all() {
    release();
    purge();
  }
};

template <typename T>
struct Level {
  Grid g;
  int CI = 1, CJ = 1, CK = 1;      // factor from the finer level (1 or 2)
  T *c = nullptr, *w = nullptr, *s = nullptr, *b = nullptr;
  T *r = nullptr;                  // right-hand side of this level
  T *e[3] = {nullptr, nullptr, nullptr};
  uint8_t *aflag = nullptr;        // NULL: all active
  // half-precision copy scale16*(c, w, s, b) for the smoother (FP32 levels,
  // built the first time a solve asks for it; NULL before).  The sweeps on
  // it compute corrections scaled by 1/scale.
  _Float16 *c16 = nullptr, *w16 = nullptr, *s16 = nullptr, *b16 = nullptr;
  double scale16 = 1.0;  // the copy's scale (a power of two)
  bool use16 = false;    // this solve's smoother reads the copy
  double scale = 1.0;    // scale16 while use16, else 1
  bool h16() const { return c16 && use16; }
  // flux form (FP32 levels): d = c - sum of couplings; the stencil kernels
  // then read (d, w, s, b), and c16 holds scale * d
  T *d = nullptr;
  // K-cycle on this level: copy of the first cycle's output, s A of it,
  // the combined correction, and the scalars rho1, alpha1, a, alpha2
  T *kc = nullptr, *kv = nullptr, *kz = nullptr;
  T *krt = nullptr;  // the second Krylov step's right-hand side (fused kc_r + sweep)
  double *kst = nullptr;
  // east, north and top couplings of a non-symmetric level (the reference's
  // AC cycle on a 7-array matrix); NULL: symmetric storage (c, w, s, b)
  T *ae = nullptr, *an = nullptr, *at = nullptr;
  DMat<T> mat() const { return DMat<T>{c, w, s, b, ae, an, at}; }
  DMat<T> matflux() const { return DMat<T>{d, w, s, b, nullptr, nullptr, nullptr}; }
  DMat<_Float16> mat16() const {
    return DMat<_Float16>{c16, w16, s16, b16, nullptr, nullptr, nullptr};
  }
};

template <typename T>
struct Hier {
  std::vector<Level<T>> L;
  double *G = nullptr;  // dense coarsest operator (FP64)
  int m = 0;
  bool built = false;
  T *qf = nullptr;  // q = A p for the flexible beta (set while a solve is captured)
};

template <typename T>
struct KrylovVecs {
  T *x = nullptr, *r = nullptr, *q = nullptr, *p[2] = {nullptr, nullptr};
  T *v[2] = {nullptr, nullptr}, *s = nullptr, *t = nullptr, *rt = nullptr;
  T *z = nullptr;  // preconditioned residual (red-black Gauss-Seidel)
};

struct Graph {
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  long long key = -1;
  double fkey = 0.0;                         // a floating-point argument baked in
  const void *px = nullptr, *pb = nullptr;  // caller arrays baked into the graph
  long long used = 0;                        // last use (cache eviction order)
};

// Everything a captured AC-preconditioned solve bakes into its graph: the
// kernels chosen (K-cycle levels, fused post-sweeps, FP16 copies, the CG
// operator's storage, flexible beta), the scalars passed by value (omega,
// coarse factor) and the caller arrays.  Solves with equal keys replay one
// graph; a solve with other settings captures its own beside it, so that
// switching settings back and forth never destroys a graph in use.
struct GKey {
  int body = 0;  // 0 FP64 AC-PCG, 1 FP32 inner solve of the refinement, 2 reliable updates
  int chunk = 0, nu1 = 0, nu2 = 0, kmask = 0, fuse = 0, cg16 = 0, half = 0, flex = 0;
  double omega = 0.0, cfac = 0.0;
  const void *px = nullptr, *pb = nullptr;
  auto tie() const {
    return std::tie(body, chunk, nu1, nu2, kmask, fuse, cg16, half, flex, omega, cfac, px, pb);
  }
  bool operator<(const GKey &o) const { return tie() < o.tie(); }
};
constexpr size_t kGraphCacheMax = 12;

// the reference's AC cycle: its own levels (smallest-prime-factor
// coarsening), each with phi (e[0]), the right-hand side f (r) and the
// work vectors of that level's CG (r, p0, p1, q); level 0's e[1] holds the
// previous cycle's phi for the divergence restore
struct RefHier {
  std::vector<Level<double>> L;
  std::vector<Metric> M;
  std::vector<std::vector<double *>> K;
  bool built = false;
  Graph gr;              // one cycle of the loop
  Graph gfull;           // the full-cycle start
  double grr = 0.0, grt = 0.0;  // level tolerances the graphs were captured with
  int gms[3] = {-1, -1, -1};    // and stale counts
};

}  // namespace host
}  // namespace psb

struct psb_ctx {
  using Grid = psb::Grid;
  using DevState = psb::DevState;
  using Work = psb::Work;
  using Arena = psb::host::Arena;
  using Graph = psb::host::Graph;
  template <typename T>
  using Hier = psb::host::Hier<T>;
  template <typename T>
  using KrylovVecs = psb::host::KrylovVecs<T>;
  psb_layout lay{};
  Grid g{};
  hipStream_t stream = nullptr;
  Arena base;   // lives as long as the context
  Arena arena;   // matrix and work vectors: rebuilt by set_matrix
  Arena harena;  // the coarse levels of the hierarchy: rebuilt with it
  // matrix
  bool have_matrix = false, sym = true;
  double *A64[7] = {};   // c w s b e n t
  float *A32[4] = {};
  bool have32 = false;
  std::vector<uint8_t> off;  // host copy of the off mask (may be empty)
  uint8_t *aflag0 = nullptr;
  // geometry
  double *dxc = nullptr, *dyc = nullptr, *dzc = nullptr;
  // solver state
  DevState *st64 = nullptr, *st32 = nullptr;
  DevState *hst = nullptr;  // pinned host mirror
  double *hist = nullptr;
  int hist_cap = 0;
  std::vector<double> hist_host;  // refinement: FP64 residual per outer step
  bool hist_on_host = false;
  int hist_count = 0;
  Work work{};
  KrylovVecs<double> k64;
  KrylovVecs<float> k32;
  Hier<double> h64;
  Hier<float> h32;
  double *db = nullptr, *dx = nullptr;  // staging for host-array solves
  double *dy = nullptr;                 // scratch for psb_apply
  double *relres = nullptr;             // 2 doubles: sum r^2, sum b^2
  double ref_dt = 1.0;                  // dt of the running AC cycle (its residual)
  Graph gr_cgdiag, gr_cgsgs, gr_bicgs;
  std::map<psb::host::GKey, Graph> gcache;  // AC-preconditioned solves
  long long gclock = 0;
  double ac_omega = 0;      // omega of the running AC solve
  int coarse_min = 8;  // coarsening halves a direction while it stays >= this
  int coarse_rule = 0; // 0 all directions at once, 1 x and y before z
  // settings of the running AC solve, read while its graph is captured
  bool fuse = false;        // two post-sweeps in one pass (k_post2) where it applies
  bool cg_half = false;     // the FP32 CG operator reads the level-0 FP16 copy too
  int kmask = 0;            // K-cycle levels (bit l: level l) of the AC hierarchies
  bool flex = false;        // flexible (Polak-Ribiere) beta in the outer CG
  int flex_req = -1;        // psb_stop.ac_flexible of the running solve
  int version = 0;     // bumped by every change of matrix, geometry or off mask
  // per-kernel timing of one eager solve (psb_profile_solve_device)
  struct Probe {
    bool on = false;
    std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> rec;
    double bytes[PSB_K_COUNT] = {};
  } probe;
  psb::host::RefHier href;             // the reference's AC cycle
  psb::DevMetrics metrics{};           // metrics of an assembled matrix (device)
  uint8_t *off_dev = nullptr;          // solid cells of an assembled matrix (device)
  std::vector<double> gdx, gdy, gdz;   // host copies of the cell sizes (empty: unit)
  std::string err;
  // device address of the static scratch of psb_solve_device_static on the
  // device this context was created on (found once per device at create)
  double *static_scratch = nullptr;
  // fine-level kernels (bit masks of kLds* kinds, psb_api.hip) that run with
  // the operand staged in LDS where the rows tile a block (psb_lds.h)
  int lds32 = 2 | 8, lds64 = 1 | 2 | 4 | 8 | 16;
  // the K-cycle's residual update fused with the next cycle's second sweep
  // (PSB_NO_KFUSE set: separate launches, for comparisons)
  bool kfuse = true;
  // the pressure's boundary conditions for the device assembly (all Neumann,
  // value 0: walls) and the boundary source it formed (NULL: none)
  psb_bnd bnd{};
  double *fbnd = nullptr;
  // the matrix is the singular Neumann operator (every active row sums to
  // zero): the AC-preconditioned CG keeps its directions free of the
  // constant null direction only then
  bool singular = true;
};

namespace psb {
namespace host {

DMat<double> mat64(const psb_ctx *c);
Metric metric(const psb_ctx *c);
std::vector<int> level_dims(int nx, int ny, int nz, int min_dim, int rule);
template <typename T>
void ensure_krylov(psb_ctx *C, KrylovVecs<T> &K, bool bicgs);
void build_hier32(psb_ctx *C);
void destroy_graph(Graph &gr);
void finish_result(psb_result *R, const DevState &h, int outer);
void gather_host(const psb_layout &L, const double *src, std::vector<double> &dst);
void scatter_host(const psb_layout &L, const std::vector<double> &src, double *dst);

// the caller's struct was filled for this library's layout (psb_stop_defaults
// / psb_cycle_defaults set struct_size): a header of another version gives
// another size, and the library would read past the caller's struct
inline bool stop_ok(const psb_stop *p) { return p->struct_size == (int)sizeof(psb_stop); }
inline bool cycle_ok(const psb_cycle *c) { return c->struct_size == (int)sizeof(psb_cycle); }

// runs f() and maps exceptions to status codes (no exception crosses the ABI)
template <typename F>
int guarded(psb_ctx *C, F f) {
  try {
    return f();
  } catch (const HipError &e) {
    if (C) C->err = hipGetErrorString(e.e);
    return e.e == hipErrorOutOfMemory ? PSB_ENOMEM : PSB_EHIP;
  } catch (const std::bad_alloc &) {
    return PSB_ENOMEM;
  } catch (const std::invalid_argument &e) {
    if (C) C->err = e.what();
    return PSB_EINVAL;
  }
}

}  // namespace host
}  // namespace psb
