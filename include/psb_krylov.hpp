/*
 * psb_krylov.hpp -- the PSI-BOIL side of libpsb.so: a Krylov class whose
 * solve() hands Linear::solve (Src/Solver/Linear/linear.h:46-56) to the
 * MI355X library through the C ABI of psb.h.
 *
 * A PSI-BOIL build adds this header (for instance as
 * Src/Solver/Linear/Krylov/gpukrylov.h, after krylov.h) and links libpsb.so;
 * a case main then passes
 *
 *     GpuKrylov solver(d);                 // in place of CG solver(d, Prec::di())
 *     Pressure pr(p, f, uvw, time, &solver, &mixed);
 *     AC multigrid(&pr);                   // every level solves on the GPU
 *
 * The reference hands one solver object to every equation level: Centered's
 * coarse-level constructor passes it down (centered.cpp:50) and AC calls it
 * on the coarse matrices (additive_cycle_functions.cpp:89-107), so one object
 * sees several grid shapes.  GpuKrylov keeps one library context per
 * (domain, shape) -- device buffers sized for that grid -- and never reads an
 * array with another shape's layout.
 *
 * Semantics follow the reference's CG::solve / BiCGS::solve (cg.cpp:10-182):
 * x is the initial guess and is overwritten (interior cells; x.exchange() at
 * exit), the stopping norm is sqrt(sum (r/dV)^2)/N/scale with dV from x's
 * domain, MinIter / MaxIter / ResRat / ResTol / stalecount gate the loop, the
 * return value is "converged", and a name prints the reference's log line.
 * The matrix is uploaded at every call: A may change between calls (the
 * pressure matrix every time step, AC's coarse matrices every cycle).
 * Library errors (a shape the library rejects, periodic couplings across the
 * domain boundary, a HIP failure) print one line and return false.
 */
#ifndef PSB_KRYLOV_HPP
#define PSB_KRYLOV_HPP

#include <map>
#include <tuple>
#include <vector>

#include "psb.h"

class GpuKrylov : public Krylov {
 public:
  /* method: PSB_CG (the reference's CG) or PSB_BICGS (BiCGS); precond:
   * PSB_PREC_DIAG (Prec::di()), PSB_PREC_SGS (in place of Prec::ic0/ic2/ic3)
   * or PSB_PREC_AC (the library's multigrid preconditioner: for CG on a
   * symmetric matrix, and for BiCGS in FP64 on any non-singular one) */
  explicit GpuKrylov(const Domain &d, int method = PSB_CG, int precond = PSB_PREC_DIAG)
      : Krylov(d) {
    psb_stop_defaults(&base_);
    base_.method = method;
    base_.precond = precond;
  }
  ~GpuKrylov() {
    for (auto &kv : ctx_) psb_destroy(kv.second);
  }
  GpuKrylov(const GpuKrylov &) = delete;
  GpuKrylov &operator=(const GpuKrylov &) = delete;

  /* the library settings beyond the reference's arguments (precision, the
   * multigrid options of PSB_PREC_AC); solve() overwrites the stopping
   * fields from its own arguments */
  psb_stop &settings() { return base_; }
  /* psb_status of the last solve, and the contexts held (one per shape) */
  int last_status() const { return last_status_; }
  int contexts() const { return (int)ctx_.size(); }

  virtual bool solve(Matrix &A, Scalar &x, Scalar &b, const MinIter &mini, const MaxIter &mi,
                     const char *name = NULL, const ResRat &rr = ResRat(),
                     const ResTol &rt = ResTol(), const real scale = 1.0,
                     const int stalecount = -1, const bool precform = true) {
    psb_ctx *C = context_for(x, b);
    if (!C) return false;
    const psb_matrix7 M = {A.c[0][0], A.w[0][0], A.e[0][0], A.s[0][0],
                           A.n[0][0], A.b[0][0], A.t[0][0]};
    int rc = psb_set_matrix(C, &M);
    if (rc != PSB_OK) return fail("psb_set_matrix", rc);
    psb_stop s = base_;
    s.min_iter = int(mini);
    s.max_iter = int(mi);
    s.res_rat = real(rr);
    s.res_tol = real(rt);
    s.scale = scale;
    s.stalecount = stalecount;
    s.precform = precform ? 1 : 0;
    psb_result R;
    rc = psb_solve(C, x[0][0], b[0][0], &s, &R);
    last_status_ = rc;
    if (rc < 0) return fail("psb_solve", rc);
    x.exchange();
    if (name != NULL)
      boil::oout << name << ", residual = " << R.res << ", ratio = " << R.res / R.res0
                 << ", iterations = " << R.iterations << boil::endl;
    return R.converged != 0;
  }

 private:
  typedef std::tuple<const Domain *, int, int, int> Key;

  bool fail(const char *what, int rc) {
    last_status_ = rc;
    boil::oout << "GpuKrylov: " << what << ": " << psb_status_string(rc) << boil::endl;
    return false;
  }

  /* the context for x's grid: created, with x's cell sizes for dV, the first
   * time this (domain, shape) is solved on; NULL (and a log line) for a
   * layout the library cannot take */
  psb_ctx *context_for(const Scalar &x, const Scalar &b) {
    const int G = x.si();
    const bool even = x.sj() == G && x.sk() == G && x.ei() == x.ni() - G - 1 &&
                      x.ej() == x.nj() - G - 1 && x.ek() == x.nk() - G - 1;
    if (!even || b.ni() != x.ni() || b.nj() != x.nj() || b.nk() != x.nk() || b.si() != G) {
      fail("layout", PSB_EINVAL);
      return NULL;
    }
    const Key key(x.domain(), x.ni(), x.nj(), x.nk());
    std::map<Key, psb_ctx *>::iterator it = ctx_.find(key);
    if (it != ctx_.end()) return it->second;
    psb_layout L;
    L.nx = x.ni() - 2 * G;
    L.ny = x.nj() - 2 * G;
    L.nz = x.nk() - 2 * G;
    L.ghost = G;
    psb_ctx *C = NULL;
    int rc = psb_create(&C, &L);
    if (rc != PSB_OK) {
      fail("psb_create", rc);
      return NULL;
    }
    /* dV(i,j,k) = dxc(i) dyc(j) dzc(k) at x's own indices, as
     * Scalar::dot_voldiv_avg evaluates it (scalar.h:317-326) */
    std::vector<double> dx, dy, dz;
    for (int i = x.si(); i <= x.ei(); i++) dx.push_back(x.dxc(i));
    for (int j = x.sj(); j <= x.ej(); j++) dy.push_back(x.dyc(j));
    for (int k = x.sk(); k <= x.ek(); k++) dz.push_back(x.dzc(k));
    rc = psb_set_geometry(C, &dx[0], &dy[0], &dz[0]);
    if (rc != PSB_OK) {
      psb_destroy(C);
      fail("psb_set_geometry", rc);
      return NULL;
    }
    ctx_[key] = C;
    return C;
  }

  std::map<Key, psb_ctx *> ctx_;
  psb_stop base_;
  int last_status_ = PSB_OK;
};

#endif
