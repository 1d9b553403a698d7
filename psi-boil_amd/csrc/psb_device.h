// psb_device.h -- device-side building blocks of the MI355X pressure solver:
// grid geometry with division by invariant integers, the 7-point stencil
// gather (4-array symmetric or 7-array storage), block partial sums for the
// two-pass reductions and the solver state the kernels share.  gfx950 only
// (wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace psb {

constexpr int kBlock = 256;      // threads per block: 4 wave64 per workgroup
constexpr int kWaves = kBlock / 64;
constexpr int kMaxLevels = 24;
constexpr int kStaleMax = 64;
constexpr int kMaxBlocks = 2048; // largest reduction grid (partials buffer)
constexpr int kMaxPartials = 8;  // partial sums per block

// n / d for 0 <= n < 2^31 with a precomputed multiplier: m = ceil(2^(31+s)/d),
// s = ceil(log2 d); floor(n m / 2^(31+s)) = floor(n / d) because the rounding
// error of m is below 2^(31+s)/d * d / 2^(31+s+s) <= 1/d (Granlund &
// Montgomery, PLDI 1994).
struct FastDiv {
  uint32_t d, m, sh;
  __device__ __forceinline__ int div(int n) const {
    return (int)(((uint64_t)(uint32_t)n * m) >> sh);
  }
};

// Interior grid, cell P = (i*ny + j)*nz + k, k fastest (the reference's
// [i][j][k] order without ghosts).  Sizes fit 31 bits.
struct Grid {
  int nx, ny, nz;
  int sI;   // ny*nz
  int N;    // nx*ny*nz
  FastDiv dI, dZ;  // division by sI and by nz
  __device__ __forceinline__ void ijk(int P, int &i, int &j, int &k) const {
    i = dI.div(P);
    const int rem = P - i * sI;
    j = dZ.div(rem);
    k = rem - j * nz;
  }
};

// Device view of a matrix.  SYM: only c, w, s, b are stored and
// e_P = w_{P+I}, n_P = s_{P+J}, t_P = b_{P+K} (zero on the east/north/top
// walls).  Otherwise e, n, t are read.
template <typename T>
struct DMat {
  const T *c, *w, *s, *b, *e, *n, *t;
};

template <typename T, int V>
struct alignas(sizeof(T) * V) VecN {
  T v[V];
};

template <typename T, int V>
__device__ __forceinline__ VecN<T, V> ldv(const T *p) {
  return *reinterpret_cast<const VecN<T, V> *>(p);
}
template <typename T, int V>
__device__ __forceinline__ void stv(T *p, const VecN<T, V> &x) {
  *reinterpret_cast<VecN<T, V> *>(p) = x;
}
// V elements of a matrix stored in M (float, double or _Float16), as T
template <typename T, typename M, int V>
__device__ __forceinline__ VecN<T, V> ldm(const M *p) {
  if constexpr (std::is_same<T, M>::value) {
    return ldv<T, V>(p);
  } else {
    const VecN<M, V> m = ldv<M, V>(p);
    VecN<T, V> r;
#pragma unroll
    for (int q = 0; q < V; q++) r.v[q] = (T)m.v[q];
    return r;
  }
}

template <typename T, int V>
__device__ __forceinline__ VecN<T, V> zerov() {
  VecN<T, V> r;
#pragma unroll
  for (int q = 0; q < V; q++) r.v[q] = T(0);
  return r;
}

// Shared solver state: Krylov scalars and stopping data.  One instance per
// precision per context, in device memory; written only by the one-block
// finalize kernels, read by the streaming kernels that follow.
struct DevState {
  double rho, alpha, beta, omega, pq, rv;      // Krylov scalars
  double res0, res, rsq0, rsq;                 // norms
  double res_rat, res_tol, scale, ntot;        // stopping parameters
  int mini, maxit, stalecount, norm;           // norm: 0 reference, 1 L2
  int iter, done, converged, reason;           // progress (reason as oracle)
  int report_iters;
  int keepmean;  // 1: a non-singular operator (Dirichlet sides, a time term): the
                 // preconditioned residual and the inner right-hand side keep
                 // their mean; 0: the singular Neumann operator, whose constant
                 // null direction they lose
  double *hist;                                // residual per iteration
  int hist_cap, pad1;
  double stale[kStaleMax];
  double mean[kMaxLevels];                     // mean of the inner right-hand side
  double rv_ss, rv_ts, rv_tt;                  // BiCGS dots
  double zmean;                                // mean of the preconditioned residual
  // reliable updates (FP32 CG inside FP64 refinement without restarts)
  int update, outer;                           // refresh pending; refreshes done
  double rlast, delta;                         // true residual at the last refresh
  double res_true;                             // last true (FP64) residual
  double rl2_true, bsq;                        // its L2 norm; ||b||^2 (reliable updates)
};

// Whether a kernel of the preconditioner on a level of n cells returns at
// once because the solve has stopped.  Below kDoneCheckMin cells it never
// reads the flag: the load is one more round trip to memory before the
// operands on levels whose launches last a few microseconds, and running such
// a launch after the stop only rewrites that level's work vectors, which
// every later use writes before it reads them.
constexpr int kDoneCheckMin = 1 << 19;
__device__ __forceinline__ bool skip_done(const DevState *st, int n) {
  return n >= kDoneCheckMin && st->done;
}

// ---------------------------------------------------------------------------
// reductions, pass 1: wave64 butterfly, LDS across the 4 waves, one partial
// per block.  Pass 2 is a one-block kernel (k_finalize) that sums the
// partials in a fixed order.  Two passes avoid the device-scope fence a
// last-block-done scheme needs, which on MI355X writes back the XCD's L2.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int NR>
__device__ __forceinline__ void block_sum(double (&v)[NR], double (*sm)[NR]) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < NR; r++) v[r] = wave_sum(v[r]);
  if (lane == 0) {
#pragma unroll
    for (int r = 0; r < NR; r++) sm[wid][r] = v[r];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NR; r++) {
    double t = 0.0;
    for (int q = 0; q < kWaves; q++) t += sm[q][r];
    v[r] = t;
  }
  __syncthreads();
}

template <int NR>
__device__ __forceinline__ void block_partial(double (&v)[NR], double *partial) {
  __shared__ double sm[kWaves][NR];
  block_sum<NR>(v, sm);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; r++) partial[(size_t)blockIdx.x * NR + r] = v[r];
  }
}

// Block -> work mapping.  Consecutive block ids land on different XCDs
// (round-robin over the 8 XCDs of MI355X); remapping gives each XCD a
// contiguous slab of planes so the i-1/i+1 neighbours of a tile sit in the
// same L2.
__device__ __forceinline__ int xcd_block() {
  int nb = gridDim.x, b = blockIdx.x;
  if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
  return b;
}

// the blocks a streaming kernel body works with: block bl of nb (the whole
// grid for one system; one system's segment of a batched grid)
struct Span {
  int bl, nb;
};
__device__ __forceinline__ Span grid_span() { return Span{xcd_block(), (int)gridDim.x}; }

// Batched launches: one grid covers many systems.  Segment q owns blocks
// [b0, b0 + nb) of the grid (b0 and nb multiples of 8, so the XCD remap
// applies inside the segment); bmap[blockIdx.x] names the segment.
struct BSeg {
  int sys, b0, nb, pad;
};
__device__ __forceinline__ Span seg_span(const BSeg &s) {
  int b = (int)blockIdx.x - s.b0;
  if ((s.nb & 7) == 0) b = (b & 7) * (s.nb >> 3) + (b >> 3);
  return Span{b, s.nb};
}

// ---------------------------------------------------------------------------
// 7-point stencil gather.  For cell vector P (V consecutive k), returns
//   cx = c*x_P,  s1 = aW xW + aE xE,  s2 = aS xS + aN xN,  s3 = aB xB + aT xT
// so that (A x)_P = ((cx - s1) - s2) - s3 and (b - A x)_P =
// (((b - cx) + s1) + s2) + s3 round exactly like the reference's four passes
// (Src/Field/Scalar/scalar_operators.cpp:22-72).  Neighbours outside the
// domain contribute 0 (their coefficient is 0 on a Neumann wall).
// X provides vec(P, i, j, k) and at(Q, i, j, k) for the operand (a plain
// array or a value formed on the fly, e.g. z + beta p); the coordinates are
// those of the first cell addressed.
// ---------------------------------------------------------------------------
template <typename T, int V>
struct Gath {
  VecN<T, V> x, c, cx, s1, s2, s3;
};

// FLUX: A.c holds d = c - (sum of the six couplings) (zero on a Neumann
// row of the pressure matrix) and the terms are formed as
//   cx = d x_P,  s1 = aW (xW - x_P) + aE (xE - x_P), ...
// so that (A x)_P = d x_P + sum a (x_P - x_nb): for smooth x no large terms
// cancel, which keeps FP32 accurate on stiff (stretched, high-contrast)
// systems.  G.c is then d + sum a, the divisor of the Jacobi sweep.
// the operand at a j neighbour of the cell vector (a marching or LDS
// operand overloads it: there the own column lives in registers)
template <class X>
__device__ __forceinline__ auto x_nbj(const X &x, int P, int ii, int jj, int kk) {
  return x.vec(P, ii, jj, kk);
}
// ... and at an i neighbour (up: i + 1)
template <class X>
__device__ __forceinline__ auto x_nbi(const X &x, bool, int P, int ii, int jj, int kk) {
  return x.vec(P, ii, jj, kk);
}

template <typename T, int V>
__device__ __forceinline__ VecN<T, V> vsel(bool keep, const VecN<T, V> &a) {
  VecN<T, V> o;
#pragma unroll
  for (int q = 0; q < V; q++) o.v[q] = keep ? a.v[q] : T(0);
  return o;
}

// the active flags of the N cells from Q as one word, read without a
// branch: with no flag array (every cell active) the word is read from
// `any` (an array of the same cells) and replaced by all-active
template <int N, typename A>
__device__ __forceinline__ uint32_t flag_word(const uint8_t *af, int Q, const A *any) {
  static_assert(N == 1 || N == 2 || N == 4, "a vector's flags are read as one word");
  const uint8_t *p = af ? af + Q : reinterpret_cast<const uint8_t *>(any + Q);
  uint32_t w;
  if constexpr (N == 4) w = *reinterpret_cast<const uint32_t *>(p);
  else if constexpr (N == 2) w = *reinterpret_cast<const uint16_t *>(p);
  else w = *p;
  return af ? w : 0xffffffffu;
}
__device__ __forceinline__ bool flag_on(uint32_t w, int u) { return ((w >> (8 * u)) & 0xffu) != 0; }
// v where on, +0 elsewhere, by a bit mask: the compiler keeps it a select
// (a plain conditional it may turn into a branch around the work feeding v,
// and the loads behind it then wait for the branch)
template <typename T>
__device__ __forceinline__ T keep_if(bool on, T v) {
  using U = typename std::conditional<sizeof(T) == 8, uint64_t, uint32_t>::type;
  const U m = U(0) - U(on);
  return __builtin_bit_cast(T, __builtin_bit_cast(U, v) & m);
}

// A cell vector's coefficients as stored (M), loaded apart from the
// arithmetic so that a kernel can issue them early (before a barrier, or a
// plane ahead).  SYM: e = A.w at i + 1, n = A.s at j + 1, bt = A.b at k + V;
// otherwise e, n, t are the arrays themselves.  Every load is issued
// whatever the position: a neighbour outside the domain is read at the cell
// itself (an address in range) and its value replaced by 0 when the
// coefficients are used.  Without branches around them, a plane's loads
// leave together instead of in dependent rounds, each waiting for the one
// before.
template <typename M, int V, bool SYM>
struct Coef {
  VecN<M, V> c, w, s, b, e, n, t;
  M bt;
};
template <typename M, int V, bool SYM>
__device__ __forceinline__ Coef<M, V, SYM> load_coef(const Grid &g, const DMat<M> &A, int P,
                                                     int i, int j, int k) {
  Coef<M, V, SYM> C;
  C.c = ldv<M, V>(A.c + P);
  C.w = ldv<M, V>(A.w + P);
  C.s = ldv<M, V>(A.s + P);
  C.b = ldv<M, V>(A.b + P);
  if constexpr (SYM) {
    C.e = ldv<M, V>(A.w + (i < g.nx - 1 ? P + g.sI : P));
    C.n = ldv<M, V>(A.s + (j < g.ny - 1 ? P + g.nz : P));
    C.bt = A.b[(k + V) < g.nz ? P + V : P];
  } else {
    C.e = ldv<M, V>(A.e + P);
    C.n = ldv<M, V>(A.n + P);
    C.t = ldv<M, V>(A.t + P);
  }
  return C;
}

template <typename T, typename M, int V>
__device__ __forceinline__ VecN<T, V> cvtv(const VecN<M, V> &m) {
  VecN<T, V> r;
#pragma unroll
  for (int q = 0; q < V; q++) r.v[q] = (T)m.v[q];
  return r;
}

// the gather with the coefficients already loaded
template <typename T, int V, bool SYM, bool FLUX = false, class X, typename M>
__device__ __forceinline__ Gath<T, V> gather_c(const Grid &g, const Coef<M, V, SYM> &C,
                                               const X &X_, int P, int i, int j, int k) {
  Gath<T, V> G;
  const bool hw = i > 0, he = i < g.nx - 1, hs = j > 0, hn = j < g.ny - 1;
  const bool hb = k > 0, ht = (k + V) < g.nz;
  const VecN<T, V> aw = cvtv<T, M, V>(C.w), as = cvtv<T, M, V>(C.s), ab = cvtv<T, M, V>(C.b);
  VecN<T, V> ae, an, at;
  G.c = cvtv<T, M, V>(C.c);
  if constexpr (SYM) {
    ae = vsel<T, V>(he, cvtv<T, M, V>(C.e));
    an = vsel<T, V>(hn, cvtv<T, M, V>(C.n));
    const T btV = ht ? (T)C.bt : T(0);
#pragma unroll
    for (int q = 0; q < V; q++) at.v[q] = (q < V - 1) ? ab.v[q + 1] : btV;
  } else {
    ae = cvtv<T, M, V>(C.e);
    an = cvtv<T, M, V>(C.n);
    at = cvtv<T, M, V>(C.t);
  }
  G.x = X_.vec(P, i, j, k);
  const VecN<T, V> xw = vsel<T, V>(hw, x_nbi(X_, false, hw ? P - g.sI : P, hw ? i - 1 : i, j, k));
  const VecN<T, V> xe = vsel<T, V>(he, x_nbi(X_, true, he ? P + g.sI : P, he ? i + 1 : i, j, k));
  const VecN<T, V> xs = vsel<T, V>(hs, x_nbj(X_, hs ? P - g.nz : P, i, hs ? j - 1 : j, k));
  const VecN<T, V> xn = vsel<T, V>(hn, x_nbj(X_, hn ? P + g.nz : P, i, hn ? j + 1 : j, k));
  const T xb0r = X_.at(hb ? P - 1 : P, i, j, hb ? k - 1 : k);
  const T xtVr = X_.at(ht ? P + V : P, i, j, ht ? k + V : k);
  const T xb0 = hb ? xb0r : T(0), xtV = ht ? xtVr : T(0);
#pragma unroll
  for (int q = 0; q < V; q++) {
    T xb = (q == 0) ? xb0 : G.x.v[q - 1];
    T xt = (q == V - 1) ? xtV : G.x.v[q + 1];
    if constexpr (FLUX) {
      const T x = G.x.v[q];
      G.cx.v[q] = G.c.v[q] * x;
      G.s1.v[q] = aw.v[q] * (xw.v[q] - x) + ae.v[q] * (xe.v[q] - x);
      G.s2.v[q] = as.v[q] * (xs.v[q] - x) + an.v[q] * (xn.v[q] - x);
      G.s3.v[q] = ab.v[q] * (xb - x) + at.v[q] * (xt - x);
      const T cc = G.c.v[q] + (((aw.v[q] + ae.v[q]) + (as.v[q] + an.v[q])) + (ab.v[q] + at.v[q]));
      G.c.v[q] = cc > T(0) ? cc : T(1);
    } else {
      G.cx.v[q] = G.c.v[q] * G.x.v[q];
      G.s1.v[q] = aw.v[q] * xw.v[q] + ae.v[q] * xe.v[q];
      G.s2.v[q] = as.v[q] * xs.v[q] + an.v[q] * xn.v[q];
      G.s3.v[q] = ab.v[q] * xb + at.v[q] * xt;
    }
  }
  return G;
}

template <typename T, int V, bool SYM, bool FLUX = false, class X, typename M>
__device__ __forceinline__ Gath<T, V> gather(const Grid &g, const DMat<M> &A,
                                             const X &X_, int P, int i, int j,
                                             int k) {
  return gather_c<T, V, SYM, FLUX>(g, load_coef<M, V, SYM>(g, A, P, i, j, k), X_, P, i, j, k);
}

template <typename T, int V>
__device__ __forceinline__ VecN<T, V> apply_A(const Gath<T, V> &G) {
  VecN<T, V> y;
#pragma unroll
  for (int q = 0; q < V; q++) {
    T v = G.cx.v[q];
    v -= G.s1.v[q];
    v -= G.s2.v[q];
    v -= G.s3.v[q];
    y.v[q] = v;
  }
  return y;
}

template <typename T, int V>
__device__ __forceinline__ VecN<T, V> resid(const VecN<T, V> &b,
                                            const Gath<T, V> &G) {
  VecN<T, V> r;
#pragma unroll
  for (int q = 0; q < V; q++) {
    T v = b.v[q] - G.cx.v[q];
    v += G.s1.v[q];
    v += G.s2.v[q];
    v += G.s3.v[q];
    r.v[q] = v;
  }
  return r;
}

// operand: a plain array
template <typename T, int V>
struct ArrX {
  const T *x;
  __device__ __forceinline__ VecN<T, V> vec(int P, int, int, int) const {
    return ldv<T, V>(x + P);
  }
  __device__ __forceinline__ T at(int Q, int, int, int) const { return x[Q]; }
};

// operand: z + beta*p with z = r * (1/c) (diagonal preconditioner,
// preconditioner_solve.cpp:4-11 with Mc = 1/c, preconditioner_form.cpp:4-13)
template <typename T, int V>
struct DiagPX {
  const T *r, *c, *p;
  T beta;
  __device__ __forceinline__ T one(T rr, T cc, T pp) const {
    T z = rr * (T(1) / cc);
    return z + beta * pp;
  }
  __device__ __forceinline__ VecN<T, V> vec(int P, int, int, int) const {
    VecN<T, V> R = ldv<T, V>(r + P), C = ldv<T, V>(c + P), Q = ldv<T, V>(p + P), o;
#pragma unroll
    for (int q = 0; q < V; q++) o.v[q] = one(R.v[q], C.v[q], Q.v[q]);
    return o;
  }
  __device__ __forceinline__ T at(int Q, int, int, int) const { return one(r[Q], c[Q], p[Q]); }
};

// operand: (z - mean z) + beta*p with z the preconditioner output, the mean
// taken over the active cells (af == NULL: all cells active)
template <typename T, int V>
struct ZPX {
  const T *z, *p;
  const uint8_t *af;
  T beta, zmean;
  __device__ __forceinline__ T one(T zz, T pp, bool act) const {
    return (act ? zz - zmean : zz) + beta * pp;
  }
  // the flags read without a branch (flag_word)
  __device__ __forceinline__ VecN<T, V> vec(int P, int, int, int) const {
    VecN<T, V> Z = ldv<T, V>(z + P), Q = ldv<T, V>(p + P), o;
    const uint32_t act = flag_word<V>(af, P, z);
#pragma unroll
    for (int q = 0; q < V; q++) o.v[q] = one(Z.v[q], Q.v[q], flag_on(act, q));
    return o;
  }
  __device__ __forceinline__ T at(int Q, int, int, int) const {
    return one(z[Q], p[Q], flag_on(flag_word<1>(af, Q, z), 0));
  }
};

// operand: e + cfac * eH[parent] (piecewise-constant prolongation,
// additive_interpolation.cpp:12-64, scaled by the coarse-correction
// factor); solid cells (af == 0) take no correction.  The coarsening factor
// per direction is 1 or 2 (shift sI/sJ/sK = 0 or 1).
template <typename T, int V>
struct ProlX {
  const T *e, *eH;
  const uint8_t *af;
  T cfac;
  int sI, sJ, sK, nYH, nZH;
  __device__ __forceinline__ int parent(int i, int j, int k) const {
    return ((i >> sI) * nYH + (j >> sJ)) * nZH + (k >> sK);
  }
  // the loads are issued unconditionally (no branch for a missing flag
  // array, flag_word, or for the coarsening along k)
  __device__ __forceinline__ T at(int Q, int i, int j, int k) const {
    const T ev = e[Q], hv = eH[parent(i, j, k)];
    return flag_on(flag_word<1>(af, Q, e), 0) ? ev + cfac * hv : ev;
  }
  // one vector load of e, two of the V (sK = 0) or V / 2 (sK = 1) parents
  // and one of the V flags: k is a multiple of V and so are nz and the
  // coarse row length
  static_assert(V == 1 || V == 2 || V == 4, "flags of a vector are read as one word");
  __device__ __forceinline__ VecN<T, V> vec(int P, int i, int j, int k) const {
    VecN<T, V> E = ldv<T, V>(e + P), o;
    const int base = ((i >> sI) * nYH + (j >> sJ)) * nZH;
    T h[V];
    if constexpr (V >= 2) {
      // halves at k / 2 (twice) for sK = 1, at k and k + V / 2 for sK = 0
      const VecN<T, V / 2> H0 = ldv<T, V / 2>(eH + base + (k >> sK));
      const VecN<T, V / 2> H1 = ldv<T, V / 2>(eH + base + (sK ? (k >> 1) : k + V / 2));
#pragma unroll
      for (int q = 0; q < V; q++)
        h[q] = sK ? H0.v[q >> 1] : (q < V / 2 ? H0.v[q] : H1.v[q - V / 2]);
    } else {
      h[0] = eH[base + (k >> sK)];
    }
    const uint32_t act = flag_word<V>(af, P, e);
#pragma unroll
    for (int q = 0; q < V; q++) o.v[q] = flag_on(act, q) ? E.v[q] + cfac * h[q] : E.v[q];
    return o;
  }
};

}  // namespace psb
