// ref_driver.cpp -- drives the reference's own classes (Niceno/PSI-BOIL,
// compiled from its sources by oracle/ref.mk) on the benchmark's synthetic
// cases and writes what they compute, so the C oracle and the HIP library can
// be checked against the reference itself.
//
// TEST INFRASTRUCTURE ONLY.  Output files (raw little-endian float64 / int32)
// go to the directory named on the command line; tests/golden/make_golden.py
// turns small ones into committed fixtures.
//
//   ref_driver case  OUT nx ny nz nsph rmin rmax rho_in rho_out seed_sph seed_vel [cx cy cz]
//   ref_driver case-s OUT nx ny nz nsph rmin rmax rho_in rho_out seed_sph seed_vel
//              d1x dnx d1y dny d1z dnz     (the same on a non-uniform grid: the
//              reference's Grid1D stretched from end cell sizes D1 to DN,
//              given in units of h = 1 / max(nx, ny, nz))
//   ref_driver cycles OUT nx ny nz nsph rmin rmax rho_in rho_out seed_sph seed_vel [cx cy cz]
//   ref_driver cycles-bicgs OUT nx ny nz nsph rmin rmax rho_in rho_out seed_sph seed_vel
//   ref_driver cycles-nonsym OUT nx ny nz nsph rmin rmax rho_in rho_out seed_sph seed_vel
//   ref_driver case-bnd OUT nx ny nz nsph rmin rmax rho_in rho_out seed_sph seed_vel t0 v0 .. t5 v5
//   ref_driver ghost[-s] OUT nx ny nz nsph rmin rmax rho_in rho_out seed_sph seed_vel sigma seed [stretch]
//   ref_driver bck OUT IN.bck nx ny nz   (Scalar::load of IN, Scalar::save of a seeded field)
//   ref_driver bench-cg nx nsph rmin rmax rho_in rho_out seed_sph seed_vel maxit
//   ref_driver bench-ac nx nsph rmin rmax rho_in rho_out seed_sph seed_vel cycles
//   (bench modes: optional cx cy cz after the limit for one centred sphere)
//
// Private and protected members of the reference's classes (AC::L,
// Centered::A, ...) are reached by widening access before its headers are
// read; the layout of the classes does not change.
#include <algorithm>
#include <array>
#include <cassert>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <list>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#define private public
#define protected public
#include "Domain/domain.h"
#include "Equation/Centered/Pressure/pressure.h"
#include "Field/Scalar/scalar.h"
#include "Field/Vector/vector.h"
#include "Grid/grid1d.h"
#include "Matter/matter.h"
#include "SimulationTime/simulation_time.h"
#include "Solver/Additive/additive.h"
#include "Solver/Linear/Iterative/iterative.h"
#include "Solver/Linear/Krylov/krylov.h"
#undef private
#undef protected

extern "C" {
#include "psb_oracle.h"
}

namespace {

std::string outdir;

void dump(const std::string &name, const std::vector<double> &v) {
  std::ofstream f(outdir + "/" + name + ".f64", std::ios::binary);
  f.write(reinterpret_cast<const char *>(v.data()), v.size() * sizeof(double));
}

void dumpi(const std::string &name, const std::vector<int> &v) {
  std::ofstream f(outdir + "/" + name + ".i32", std::ios::binary);
  f.write(reinterpret_cast<const char *>(v.data()), v.size() * sizeof(int));
}

// interior of a Scalar in [i][j][k] order
std::vector<double> interior(const Scalar &s) {
  std::vector<double> v;
  v.reserve((size_t)(s.ei() - s.si() + 1) * (s.ej() - s.sj() + 1) * (s.ek() - s.sk() + 1));
  for (int i = s.si(); i <= s.ei(); i++)
    for (int j = s.sj(); j <= s.ej(); j++)
      for (int k = s.sk(); k <= s.ek(); k++) v.push_back(s[i][j][k]);
  return v;
}

std::vector<int> interior_i(const ScalarInt &s) {
  std::vector<int> v;
  for (int i = s.si(); i <= s.ei(); i++)
    for (int j = s.sj(); j <= s.ej(); j++)
      for (int k = s.sk(); k <= s.ek(); k++) v.push_back(s[i][j][k]);
  return v;
}

void set_interior(Scalar &s, const std::vector<double> &v) {
  size_t q = 0;
  for (int i = s.si(); i <= s.ei(); i++)
    for (int j = s.sj(); j <= s.ej(); j++)
      for (int k = s.sk(); k <= s.ek(); k++) s[i][j][k] = v[q++];
}

void neumann(Scalar &s) {
  s.bc().add(BndCnd(Dir::imin(), BndType::neumann()));
  s.bc().add(BndCnd(Dir::imax(), BndType::neumann()));
  s.bc().add(BndCnd(Dir::jmin(), BndType::neumann()));
  s.bc().add(BndCnd(Dir::jmax(), BndType::neumann()));
  s.bc().add(BndCnd(Dir::kmin(), BndType::neumann()));
  s.bc().add(BndCnd(Dir::kmax(), BndType::neumann()));
}

std::vector<double> seeded(uint64_t seed, size_t n) {
  std::vector<double> v(n);
  uint64_t st = seed;
  for (auto &x : v) x = 2.0 * ora_u01(&st) - 1.0;
  return v;
}

// Pressure whose right-hand side is the reference's update_rhs() followed by
// the zero-mean shift the benchmark cases add.
class CasePressure : public Pressure {
 public:
  CasePressure(const Scalar &p, const Scalar &f, const Vector &u, Times &t,
               Linear *sm, Matter *m)
      : Pressure(p, f, u, t, sm, m) {}
  real update_rhs() {
    real r = Pressure::update_rhs();
    double sum = 0.0;
    long cnt = 0;
    for_ijk(i, j, k) { sum += fnew[i][j][k]; cnt++; }
    double mean = sum / (double)cnt;
    for_ijk(i, j, k) fnew[i][j][k] -= mean;
    return r;
  }
};

struct Case {
  int nx, ny, nz, nsph;
  double rmin, rmax, rho_in, rho_out;
  uint64_t seed_sph, seed_vel;
  bool centred = false;
  double cx = 0, cy = 0, cz = 0;
  // first and last cell size per direction in units of h (1, 1: uniform);
  // Grid1D distributes the nodes in between (grid1d_distribute_nodes_inside.cpp)
  double d1[3] = {1.0, 1.0, 1.0}, dn[3] = {1.0, 1.0, 1.0};
  // pressure boundary condition per side (imin, imax, jmin, jmax, kmin, kmax):
  // 0 neumann, 1 dirichlet, with its value; pbnd false: neumann 0 everywhere
  bool pbnd = false;
  int pbt[6] = {0, 0, 0, 0, 0, 0};
  double pbv[6] = {0, 0, 0, 0, 0, 0};
};

struct World {
  Grid1D *gx, *gy, *gz;
  Domain *d;
  Times *time;
  Krylov *solver;
  Vector *uvw;
  Scalar *p, *f, *c;
  Matter *in, *out, *mixed;
  CasePressure *pr;
  std::vector<double> sph;
};

// make_solver: the solver every level of the case's Pressure holds (NULL: the
// reference's CG with Prec::di())
World build_world(const Case &cs, Krylov *(*make_solver)(const Domain &) = nullptr) {
  World w;
  const int nmax = std::max(cs.nx, std::max(cs.ny, cs.nz));
  const double h = 1.0 / nmax;
  w.gx = new Grid1D(Range<real>(0.0, cs.nx * h), Range<real>(cs.d1[0] * h, cs.dn[0] * h), cs.nx,
                    Periodic::no());
  w.gy = new Grid1D(Range<real>(0.0, cs.ny * h), Range<real>(cs.d1[1] * h, cs.dn[1] * h), cs.ny,
                    Periodic::no());
  w.gz = new Grid1D(Range<real>(0.0, cs.nz * h), Range<real>(cs.d1[2] * h, cs.dn[2] * h), cs.nz,
                    Periodic::no());
  w.d = new Domain(*w.gx, *w.gy, *w.gz);
  w.time = new Times(1, 1.0);
  w.solver = make_solver ? make_solver(*w.d) : new CG(*w.d, Prec::di());
  w.uvw = new Vector(*w.d);
  w.p = new Scalar(*w.d);
  w.f = new Scalar(*w.d);
  w.c = new Scalar(*w.d);
  for_m(m) {
    w.uvw->bc(m).add(BndCnd(Dir::imin(), BndType::wall()));
    w.uvw->bc(m).add(BndCnd(Dir::imax(), BndType::wall()));
    w.uvw->bc(m).add(BndCnd(Dir::jmin(), BndType::wall()));
    w.uvw->bc(m).add(BndCnd(Dir::jmax(), BndType::wall()));
    w.uvw->bc(m).add(BndCnd(Dir::kmin(), BndType::wall()));
    w.uvw->bc(m).add(BndCnd(Dir::kmax(), BndType::wall()));
  }
  neumann(*w.c);
  if (!cs.pbnd) {
    neumann(*w.p);
  } else {
    const Dir dirs[6] = {Dir::imin(), Dir::imax(), Dir::jmin(), Dir::jmax(), Dir::kmin(),
                         Dir::kmax()};
    for (int q = 0; q < 6; q++)
      w.p->bc().add(BndCnd(dirs[q], cs.pbt[q] ? BndType::dirichlet() : BndType::neumann(),
                           cs.pbv[q]));
  }
  // colour function: 1 inside any sphere (cell-centre test), 0 outside
  if (cs.centred) {
    w.sph = {cs.cx, cs.cy, cs.cz, cs.rmin};
  } else {
    w.sph.assign(4 * std::max(cs.nsph, 1), 0.0);
    ora_gen_spheres(cs.seed_sph, cs.nsph, cs.rmin, cs.rmax, cs.nx * h, cs.ny * h,
                    cs.nz * h, w.sph.data());
    w.sph.resize(4 * cs.nsph);
  }
  Scalar &c = *w.c;
  for_vijk(c, i, j, k) {
    bool inside = false;
    for (size_t s = 0; s + 3 < w.sph.size() && !inside; s += 4) {
      double dx = c.xc(i) - w.sph[s], dy = c.yc(j) - w.sph[s + 1],
             dz = c.zc(k) - w.sph[s + 2];
      if (std::sqrt(dx * dx + dy * dy + dz * dz) < w.sph[s + 3]) inside = true;
    }
    c[i][j][k] = inside ? 1.0 : 0.0;
  }
  c.bnd_update();
  c.exchange();
  w.in = new Matter(*w.d);
  w.out = new Matter(*w.d);
  w.in->rho(cs.rho_in);
  w.in->mu(1.0e-3);
  w.out->rho(cs.rho_out);
  w.out->mu(1.0e-5);
  w.mixed = new Matter(*w.in, *w.out, w.c);
  // face velocities from the shared generator, wall-normal faces 0
  std::vector<double> u((size_t)(cs.nx + 1) * cs.ny * cs.nz),
      v((size_t)cs.nx * (cs.ny + 1) * cs.nz), ww((size_t)cs.nx * cs.ny * (cs.nz + 1));
  ora_face_velocity(cs.seed_vel, cs.nx, cs.ny, cs.nz, u.data(), v.data(), ww.data());
  Vector &U = *w.uvw;
  const int si = w.p->si(), sj = w.p->sj(), sk = w.p->sk();
  for (int i = 0; i <= cs.nx; i++)
    for (int j = 0; j < cs.ny; j++)
      for (int k = 0; k < cs.nz; k++)
        U[Comp::u()][si + i][sj + j][sk + k] = u[((size_t)i * cs.ny + j) * cs.nz + k];
  for (int i = 0; i < cs.nx; i++)
    for (int j = 0; j <= cs.ny; j++)
      for (int k = 0; k < cs.nz; k++)
        U[Comp::v()][si + i][sj + j][sk + k] = v[((size_t)i * (cs.ny + 1) + j) * cs.nz + k];
  for (int i = 0; i < cs.nx; i++)
    for (int j = 0; j < cs.ny; j++)
      for (int k = 0; k <= cs.nz; k++)
        U[Comp::w()][si + i][sj + j][sk + k] = ww[((size_t)i * cs.ny + j) * (cs.nz + 1) + k];
  w.pr = new CasePressure(*w.p, *w.f, *w.uvw, *w.time, w.solver, w.mixed);
  w.pr->update_rhs();
  return w;
}

void dump_case(World &w, const Case &cs) {
  CasePressure &pr = *w.pr;
  const Domain &d = *w.d;
  const Scalar &p = pr.phi;
  std::vector<double> xn, yn, zn, m;
  for (int i = p.si(); i <= p.ei() + 1; i++) xn.push_back(d.xn(i));
  for (int j = p.sj(); j <= p.ej() + 1; j++) yn.push_back(d.yn(j));
  for (int k = p.sk(); k <= p.ek() + 1; k++) zn.push_back(d.zn(k));
  dump("xn", xn); dump("yn", yn); dump("zn", zn);
  // metrics as Pressure::discretize reads them
  std::vector<double> dxc, dxw, dxe, dyc, dys, dyn, dzc, dzb, dzt;
  for (int i = p.si(); i <= p.ei(); i++) { dxc.push_back(d.dxc(i)); dxw.push_back(d.dxw(i)); dxe.push_back(d.dxe(i)); }
  for (int j = p.sj(); j <= p.ej(); j++) { dyc.push_back(d.dyc(j)); dys.push_back(d.dys(j)); dyn.push_back(d.dyn(j)); }
  for (int k = p.sk(); k <= p.ek(); k++) { dzc.push_back(d.dzc(k)); dzb.push_back(d.dzb(k)); dzt.push_back(d.dzt(k)); }
  dump("dxc", dxc); dump("dxw", dxw); dump("dxe", dxe);
  dump("dyc", dyc); dump("dys", dys); dump("dyn", dyn);
  dump("dzc", dzc); dump("dzb", dzb); dump("dzt", dzt);
  dump("spheres", w.sph);
  // cell density with a one-cell ghost layer, as Property reads it
  std::vector<double> rho;
  for (int i = p.si() - 1; i <= p.ei() + 1; i++)
    for (int j = p.sj() - 1; j <= p.ej() + 1; j++)
      for (int k = p.sk() - 1; k <= p.ek() + 1; k++) rho.push_back(w.mixed->rho(i, j, k));
  dump("rho", rho);
  Matrix &A = pr.A;
  dump("A_c", interior(A.c)); dump("A_w", interior(A.w)); dump("A_e", interior(A.e));
  dump("A_s", interior(A.s)); dump("A_n", interior(A.n)); dump("A_b", interior(A.b));
  dump("A_t", interior(A.t)); dump("A_ci", interior(A.ci));
  std::vector<double> dV;
  for_vijk(p, i, j, k) dV.push_back(p.dV(i, j, k));
  dump("dV", dV);
  dump("b", interior(pr.fnew));
  // operators on a seeded vector
  const size_t N = (size_t)cs.nx * cs.ny * cs.nz;
  std::vector<double> xr = seeded(7, N);
  Scalar x(*w.d), y(*w.d);
  neumann(x);
  set_interior(x, xr);
  dump("x_rand", xr);
  y = A * x;
  dump("Ax_rand", interior(y));
  y = pr.fnew - A * x;
  dump("r_rand", interior(y));
  std::cout << "DOT " << x.dot(pr.fnew) << " " << x.dot_voldiv_avg(pr.fnew) << "\n";
}

void run_krylov(World &w, Krylov *sol, const char *tag, const std::vector<int> &ks,
                int maxit_full, double rr) {
  CasePressure &pr = *w.pr;
  for (int k : ks) {
    Scalar x(*w.d);
    neumann(x);
    x = 0.0;
    std::cout << "RUN " << tag << " " << k << "\n";
    bool ok = sol->solve(pr.A, x, pr.fnew, MinIter(1), MaxIter(k), tag, ResRat(rr),
                         ResTol(boil::femto), 1.0, -1, true);
    std::cout << "RET " << tag << " " << k << " " << ok << "\n";
    dump(std::string(tag) + "_x" + std::to_string(k), interior(x));
  }
  Scalar x(*w.d);
  neumann(x);
  x = 0.0;
  std::cout << "RUN " << tag << " full\n";
  bool ok = sol->solve(pr.A, x, pr.fnew, MinIter(1), MaxIter(maxit_full), tag,
                       ResRat(rr), ResTol(boil::femto), 1.0, -1, true);
  std::cout << "RET " << tag << " full " << ok << "\n";
  dump(std::string(tag) + "_xfull", interior(x));
}

void dump_ac(World &w) {
  CasePressure &pr = *w.pr;
  AC ac(&pr);
  std::cout << "NLEV " << ac.nlevels << "\n";
  for (int l = 1; l < ac.nlevels; l++) ac.coarsen_system(*ac.L[l - 1], *ac.L[l]);
  for (int l = 0; l < ac.nlevels; l++) {
    Centered &L = *ac.L[l];
    std::cout << "LEVEL " << l << " " << L.ei() - L.si() + 1 << " " << L.ej() - L.sj() + 1
              << " " << L.ek() - L.sk() + 1 << "\n";
    std::string s = "L" + std::to_string(l) + "_";
    if (l > 0) {
      dump(s + "c", interior(L.A.c)); dump(s + "w", interior(L.A.w));
      dump(s + "e", interior(L.A.e)); dump(s + "s", interior(L.A.s));
      dump(s + "n", interior(L.A.n)); dump(s + "b", interior(L.A.b));
      dump(s + "t", interior(L.A.t));
    }
    dumpi(s + "aflag", interior_i(L.aflag));
  }
  // restriction of level 0 -> 1 with a seeded phi
  if (ac.nlevels > 1) {
    Centered &h = *ac.L[0];
    Centered &H = *ac.L[1];
    std::vector<double> ph = seeded(11, interior(h.phi).size());
    std::vector<double> ph0 = interior(h.phi);
    set_interior(h.phi, ph);
    ac.restriction(h, H);
    dump("restrict_phi", ph);
    dump("restrict_fH", interior(H.fnew));
    set_interior(h.phi, ph0);
    std::vector<double> pH = seeded(13, interior(H.phi).size());
    set_interior(H.phi, pH);
    h.phi = 0.0;
    ac.interpolation(H, h);
    dump("interp_phiH", pH);
    dump("interp_phih", interior(h.phi));
    H.phi = 0.0;
    H.fnew = 0.0;
    h.phi = 0.0;
  }
  // V-cycles: phi after c cycles from phi = 0, and a converged history
  for (int c = 1; c <= 3; c++) {
    pr.phi = 0.0;
    ac.max_cycles(c);
    std::cout << "ACRUN " << c << "\n";
    int ncyc = 0;
    bool ok = ac.vcycle(ResRat(1e-30), &ncyc);
    std::cout << "ACRET " << c << " " << ok << " " << ncyc << "\n";
    dump("ac_phi" + std::to_string(c), interior(pr.phi));
  }
  pr.phi = 0.0;
  ac.max_cycles(20);
  std::cout << "ACRUN full\n";
  int ncyc = 0;
  bool ok = ac.vcycle(ResRat(1e-8), &ncyc);
  std::cout << "ACRET full " << ok << " " << ncyc << "\n";
  dump("ac_phifull", interior(pr.phi));
}

// AC::cycle with the other cycle shapes and controls (additive_cycle.cpp,
// additive_recursive.cpp:37-166): each run starts from phi = 0 with the
// constructor's settings, is framed by ACRUN / ACRET lines (the cycle log
// lines in between) and leaves phi in <tag>_phi.f64
// bicgs: the world's levels hold BiCGS with Prec::di() (build_world's
// make_solver), the runs a subset of the CG ones, tags prefixed "B"
void dump_cycles(World &w, bool bicgs = false) {
  CasePressure &pr = *w.pr;
  struct Run {
    std::string tag;
    int kind;  // 0 V, 1 F, 2 W; +10: full-cycle (V) initialisation
    double rr;
    int maxc;
    bool linf, stopdiv, prio;
    int minc;
    double smooth_rr;  // AC::resrat of the level solves (atto by default)
    std::array<int, 3> mv;
    std::array<int, 3> ms;  // AC::stale_iter_array (-1: off, the default)
  };
  const double A = boil::atto;
  const std::array<int, 3> NS = {-1, -1, -1};
  const std::vector<Run> runs = {
      {"F1", 1, 1e-30, 1, false, true, true, 0, A, {20, 20, 20}, NS},
      {"F2", 1, 1e-30, 2, false, true, true, 0, A, {20, 20, 20}, NS},
      {"Ffull", 1, 1e-8, 20, false, true, true, 0, A, {20, 20, 20}, NS},
      {"W1", 2, 1e-30, 1, false, true, true, 0, A, {20, 20, 20}, NS},
      {"W2", 2, 1e-30, 2, false, true, true, 0, A, {20, 20, 20}, NS},
      {"Wfull", 2, 1e-8, 20, false, true, true, 0, A, {20, 20, 20}, NS},
      {"FMV2", 10, 1e-30, 2, false, true, true, 0, A, {20, 20, 20}, NS},
      {"FMVfull", 10, 1e-8, 20, false, true, true, 0, A, {20, 20, 20}, NS},
      {"FMF2", 11, 1e-30, 2, false, true, true, 0, A, {20, 20, 20}, NS},
      {"FMW2", 12, 1e-30, 2, false, true, true, 0, A, {20, 20, 20}, NS},
      {"Vlinf", 0, 1e-8, 20, true, true, true, 0, A, {20, 20, 20}, NS},
      {"Vdiv", 0, 1e-30, 20, false, true, true, 0, A, {20, 20, 20}, NS},
      {"Vnodiv", 0, 1e-30, 20, false, false, true, 0, A, {20, 20, 20}, NS},
      {"Vmin", 0, 1e-30, 20, false, true, true, 12, A, {20, 20, 20}, NS},
      {"Vminnoprio", 0, 1e-30, 20, false, true, false, 12, A, {20, 20, 20}, NS},
      {"Vlinfdiv", 0, 1e-30, 8, true, true, true, 0, A, {1, 1, 1}, NS},
      {"Vmv", 0, 1e-8, 20, false, true, true, 0, A, {3, 20, 5}, NS},
      {"Vsmooth", 0, 1e-8, 20, false, true, true, 0, 1e-2, {20, 20, 20}, NS},
      {"Vstale", 0, 1e-8, 20, false, true, true, 0, A, {20, 20, 20}, {3, 3, 3}},
      {"Fstale", 1, 1e-8, 20, false, true, true, 0, A, {20, 20, 20}, {2, 4, 2}},
  };
  // Cycle::flex is left out: its recursion (additive_recursive.cpp:167-200)
  // ends only when a level's smoother meets its ResRat, and once a level's
  // residual reaches round-off it alternates between two levels for good
  // (this driver never returned from it on these cases)
  const std::vector<std::string> bicgs_tags = {"F2", "W1", "FMV2", "Vmv", "Vsmooth", "Vstale"};
  for (Run r : runs) {
    if (bicgs) {
      if (std::find(bicgs_tags.begin(), bicgs_tags.end(), r.tag) == bicgs_tags.end()) continue;
      r.tag = "B" + r.tag;
    }
    AC ac(&pr);
    pr.phi = 0.0;
    ac.max_cycles(r.maxc);
    ac.min_cycles(r.minc);
    ac.use_linf_error(r.linf);
    ac.stop_if_diverging(r.stopdiv);
    ac.priority_min_cyc = r.prio;
    ac.resrat(ResRat(r.smooth_rr));
    std::array<MaxIter, 3> mv = {MaxIter(r.mv[0]), MaxIter(r.mv[1]), MaxIter(r.mv[2])};
    ac.max_iter_array(mv);
    std::array<MaxIter, 3> ms = {MaxIter(r.ms[0]), MaxIter(r.ms[1]), MaxIter(r.ms[2])};
    ac.stale_iter_array(ms);
    std::cout << "ACRUN " << r.tag << " " << r.kind << " " << r.rr << " " << r.maxc << " "
              << r.linf << " " << r.stopdiv << " " << r.prio << " " << r.minc << " "
              << r.smooth_rr << " " << r.mv[0] << " " << r.mv[1] << " " << r.mv[2] << " "
              << r.ms[0] << " " << r.ms[1] << " " << r.ms[2] << std::endl;
    int ncyc = 0;
    const Cycle loops[3] = {Cycle::V(), Cycle::F(), Cycle::W()};
    const Cycle init = r.kind >= 10 ? Cycle::V() : Cycle::none();
    bool ok = ac.cycle(init, loops[r.kind % 10], ResTol(-1.0), ResRat(r.rr), ac.mv_def,
                       ac.ms_def, &ncyc);
    std::cout << "ACRET " << r.tag << " " << ok << " " << ncyc << std::endl;
    dump(r.tag + "_phi", interior(pr.phi));
  }
}

// a non-symmetric system in the reference's own storage, as the equations
// that hold BiCGS make (EnthalpyFD's couplings differ on the two sides of a
// face near an interface, and a time term adds to the diagonal): the east
// and top couplings of each interior cell scaled by 1 + fe, 1 + ft (fixed
// patterns of cell indices in [0, 0.25]), the diagonal raised by the added
// coupling and by 5 % of itself, ci = 1 / c.  The matrix is written as
// N_A_<c,w,e,s,n,b,t>.f64.
void make_nonsymmetric(CasePressure &pr) {
  Matrix &A = pr.A;
  const Scalar &p = pr.phi;
  for (int i = p.si(); i <= p.ei(); i++)
    for (int j = p.sj(); j <= p.ej(); j++)
      for (int k = p.sk(); k <= p.ek(); k++) {
        const int ii = i - p.si(), jj = j - p.sj(), kk = k - p.sk();
        const double fe = 0.025 * ((7 * ii + 3 * jj + 5 * kk) % 11);
        const double ft = 0.25 / 6.0 * ((2 * ii + 5 * jj + 3 * kk) % 7);
        const double de = fe * A.e[i][j][k], dt = ft * A.t[i][j][k];
        A.e[i][j][k] += de;
        A.t[i][j][k] += dt;
        A.c[i][j][k] += de + dt;
        A.c[i][j][k] += 0.05 * A.c[i][j][k];
        A.ci[i][j][k] = 1.0 / A.c[i][j][k];
      }
  dump("N_A_c", interior(A.c)); dump("N_A_w", interior(A.w)); dump("N_A_e", interior(A.e));
  dump("N_A_s", interior(A.s)); dump("N_A_n", interior(A.n)); dump("N_A_b", interior(A.b));
  dump("N_A_t", interior(A.t));
}

// a field with its one-cell ghost layer, [i][j][k] with k fastest
std::vector<double> ghosted(const Scalar &s) {
  std::vector<double> v;
  for (int i = s.si() - 1; i <= s.ei() + 1; i++)
    for (int j = s.sj() - 1; j <= s.ej() + 1; j++)
      for (int k = s.sk() - 1; k <= s.ek() + 1; k++) v.push_back(s[i][j][k]);
  return v;
}

// Pressure::ghost (pressure_ghost.cpp:10-190): the ghost-fluid method's
// surface-tension source added to fext, for the case's colour function, the
// mixture's surface tension sigma and a seeded curvature field (15 % zeros,
// the rest in [-20, 30), so both of kappa_ave*'s averages occur), Neumann
// ghosts.  Writes c_g, kappa_g (ghosted) and fext (interior).
void dump_ghost(World &w, double sigma, uint64_t seed) {
  CasePressure &pr = *w.pr;
  w.mixed->sigma(sigma);
  Scalar kappa(*w.d);
  neumann(kappa);
  uint64_t st = seed;
  for_vijk(kappa, i, j, k) {
    const double u = ora_u01(&st);
    kappa[i][j][k] = u < 0.15 ? 0.0 : (u - 0.4) * 50.0;
  }
  kappa.bnd_update();
  kappa.exchange();
  pr.fext = 0.0;
  pr.ghost(*w.c, kappa);
  dump("c_g", ghosted(*w.c));
  dump("kappa_g", ghosted(kappa));
  dump("fext", interior(pr.fext));
}

// AC::flexcycle (additive.h:54-95, additive_recursive.cpp:161-200): from phi
// = 0, the level solves' ResRat given (the recursion climbs back only when a
// level's smoother meets it); tag, cycles and phi as dump_cycles writes them
void dump_flex(World &w, double smooth_rr, int maxc, const std::array<int, 3> &mvs) {
  CasePressure &pr = *w.pr;
  AC ac(&pr);
  pr.phi = 0.0;
  ac.max_cycles(maxc);
  ac.resrat(ResRat(smooth_rr));
  std::array<MaxIter, 3> mv = {MaxIter(mvs[0]), MaxIter(mvs[1]), MaxIter(mvs[2])};
  ac.max_iter_array(mv);
  std::cout << "ACRUN Flex 3 1e-08 " << maxc << " 0 1 1 0 " << smooth_rr << " " << mvs[0] << " "
            << mvs[1] << " " << mvs[2] << " -1 -1 -1" << std::endl;
  int ncyc = 0;
  bool ok = ac.flexcycle(ResRat(1e-8), &ncyc);
  std::cout << "ACRET Flex " << ok << " " << ncyc << std::endl;
  dump("Flex_phi", interior(pr.phi));
}

Case parse_case(char **a, int nx, int ny, int nz) {
  Case cs;
  cs.nx = nx; cs.ny = ny; cs.nz = nz;
  cs.nsph = std::atoi(a[0]);
  cs.rmin = std::atof(a[1]);
  cs.rmax = std::atof(a[2]);
  cs.rho_in = std::atof(a[3]);
  cs.rho_out = std::atof(a[4]);
  cs.seed_sph = std::strtoull(a[5], nullptr, 10);
  cs.seed_vel = std::strtoull(a[6], nullptr, 10);
  return cs;
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

#ifndef REF_DRIVER_EMBED
int main(int argc, char **argv) {
  std::cout << std::setprecision(17);
  if (argc < 2) {
    std::fprintf(stderr, "usage: ref_driver case|cycles|cycles-bicgs|flex|prec|bck|bench-cg|bench-ic2|bench-ac ...\n");
    return 2;
  }
  std::string mode = argv[1];
  if (mode == "case" || mode == "case-s") {
    if (argc < 13) { std::fprintf(stderr, "case: too few arguments\n"); return 2; }
    outdir = argv[2];
    Case cs = parse_case(argv + 6, std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]));
    if (mode == "case-s") {
      if (argc < 19) { std::fprintf(stderr, "case-s: too few arguments\n"); return 2; }
      for (int q = 0; q < 3; q++) {
        cs.d1[q] = std::atof(argv[13 + 2 * q]);
        cs.dn[q] = std::atof(argv[14 + 2 * q]);
      }
    } else if (argc >= 16) {
      cs.centred = true;
      cs.cx = std::atof(argv[13]); cs.cy = std::atof(argv[14]); cs.cz = std::atof(argv[15]);
    }
    World w = build_world(cs);
    dump_case(w, cs);
    run_krylov(w, w.solver, "cg", {1, 2, 3, 5, 8, 13, 21}, 2000, 1e-8);
    BiCGS bi(*w.d, Prec::di());
    run_krylov(w, &bi, "bicgs", {1, 2, 3, 5, 8}, 2000, 1e-8);
    Jacobi jac(*w.d);
    {
      Scalar x(*w.d);
      neumann(x);
      x = 0.0;
      bool ok = jac.solve(w.pr->A, x, w.pr->fnew, MinIter(1), MaxIter(5), "jacobi",
                          ResRat(1e-30), ResTol(0.0), 1.0, -1, true);
      std::cout << "RET jacobi 5 " << ok << "\n";
      dump("jacobi_x5", interior(x));
    }
    dump_ac(w);
    return 0;
  }
  if (mode == "cycles") {
    if (argc < 13) { std::fprintf(stderr, "cycles: too few arguments\n"); return 2; }
    outdir = argv[2];
    Case cs = parse_case(argv + 6, std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]));
    if (argc >= 16) {
      cs.centred = true;
      cs.cx = std::atof(argv[13]); cs.cy = std::atof(argv[14]); cs.cz = std::atof(argv[15]);
    }
    World w = build_world(cs);
    dump_cycles(w);
    return 0;
  }
  if (mode == "cycles-bicgs") {
    // ref_driver cycles-bicgs OUT nx ny nz nsph rmin rmax rho_in rho_out seed_sph seed_vel:
    // AC::cycle runs with BiCGS (Prec::di()) as every level's solver
    if (argc < 13) { std::fprintf(stderr, "cycles-bicgs: too few arguments\n"); return 2; }
    outdir = argv[2];
    Case cs = parse_case(argv + 6, std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]));
    World w = build_world(cs, [](const Domain &d) -> Krylov * { return new BiCGS(d, Prec::di()); });
    dump_cycles(w, true);
    return 0;
  }
  if (mode == "ghost" || mode == "ghost-s") {
    // ref_driver ghost OUT nx ny nz nsph rmin rmax rho_in rho_out seed_sph seed_vel sigma seed
    //            [d1x dnx d1y dny d1z dnz (ghost-s)]: the case (as "case" writes it) and
    //            Pressure::ghost's fext
    if (argc < 15) { std::fprintf(stderr, "ghost: too few arguments\n"); return 2; }
    outdir = argv[2];
    Case cs = parse_case(argv + 6, std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]));
    if (mode == "ghost-s") {
      if (argc < 21) { std::fprintf(stderr, "ghost-s: too few arguments\n"); return 2; }
      for (int q = 0; q < 3; q++) {
        cs.d1[q] = std::atof(argv[15 + 2 * q]);
        cs.dn[q] = std::atof(argv[16 + 2 * q]);
      }
    }
    World w = build_world(cs);
    dump_case(w, cs);
    dump_ghost(w, std::atof(argv[13]), std::strtoull(argv[14], nullptr, 10));
    return 0;
  }
  if (mode == "case-bnd") {
    // ref_driver case-bnd OUT nx ny nz nsph rmin rmax rho_in rho_out seed_sph seed_vel
    //                     t0 v0 .. t5 v5: the case with the pressure's boundary
    //                     condition per side (0 neumann, 1 dirichlet; value), written
    //                     as "case" writes it, plus fbnd (Centered::create_system_bnd)
    if (argc < 25) { std::fprintf(stderr, "case-bnd: too few arguments\n"); return 2; }
    outdir = argv[2];
    Case cs = parse_case(argv + 6, std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]));
    cs.pbnd = true;
    for (int q = 0; q < 6; q++) {
      cs.pbt[q] = std::atoi(argv[13 + 2 * q]);
      cs.pbv[q] = std::atof(argv[14 + 2 * q]);
    }
    World w = build_world(cs);
    dump_case(w, cs);
    dump("fbnd", interior(w.pr->fbnd));
    return 0;
  }
  if (mode == "cycles-nonsym") {
    // ref_driver cycles-nonsym OUT nx ny nz nsph rmin rmax rho_in rho_out seed_sph seed_vel:
    // the cycles-bicgs runs on the case's matrix made non-symmetric
    // (make_nonsymmetric), which AC coarsens at the start of every cycle
    if (argc < 13) { std::fprintf(stderr, "cycles-nonsym: too few arguments\n"); return 2; }
    outdir = argv[2];
    Case cs = parse_case(argv + 6, std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]));
    World w = build_world(cs, [](const Domain &d) -> Krylov * { return new BiCGS(d, Prec::di()); });
    make_nonsymmetric(*w.pr);
    dump_cycles(w, true);
    return 0;
  }
  if (mode == "flex") {
    // ref_driver flex OUT nx ny nz nsph rmin rmax rho_in rho_out seed_sph seed_vel
    //                 smooth_rr max_cycles mv0 mv1 mv2
    if (argc < 18) { std::fprintf(stderr, "flex: too few arguments\n"); return 2; }
    outdir = argv[2];
    Case cs = parse_case(argv + 6, std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]));
    World w = build_world(cs);
    dump_flex(w, std::atof(argv[13]), std::atoi(argv[14]),
              {std::atoi(argv[15]), std::atoi(argv[16]), std::atoi(argv[17])});
    return 0;
  }
  if (mode == "prec") {
    // CG with the reference's incomplete-Cholesky preconditioners
    // (Prec::ic0, Prec::ic2; linear.cpp:31-34) on the case's pressure system
    if (argc < 13) { std::fprintf(stderr, "prec: too few arguments\n"); return 2; }
    outdir = argv[2];
    Case cs = parse_case(argv + 6, std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]));
    if (argc >= 16) {
      cs.centred = true;
      cs.cx = std::atof(argv[13]); cs.cy = std::atof(argv[14]); cs.cz = std::atof(argv[15]);
    }
    World w = build_world(cs);
    CG ic0(*w.d, Prec::ic0());
    run_krylov(w, &ic0, "ic0", {1, 5}, 2000, 1e-8);
    CG ic2(*w.d, Prec::ic2());
    run_krylov(w, &ic2, "ic2", {1, 5}, 2000, 1e-8);
    return 0;
  }
  if (mode == "bck") {
    if (argc < 7) { std::fprintf(stderr, "bck: too few arguments\n"); return 2; }
    outdir = argv[2];
    Case cs;
    cs.nx = std::atoi(argv[4]); cs.ny = std::atoi(argv[5]); cs.nz = std::atoi(argv[6]);
    cs.nsph = 0; cs.rmin = cs.rmax = 0.1; cs.rho_in = cs.rho_out = 1.0;
    cs.seed_sph = cs.seed_vel = 1;
    World w = build_world(cs);
    Scalar s(*w.d);
    {
      std::ifstream in(argv[3], std::ios::binary);
      s.load(in);
    }
    dump("bck_loaded", interior(s));
    // a seeded field, ghosts included, written by the reference
    std::vector<double> full = seeded(17, (size_t)s.ni() * s.nj() * s.nk());
    std::memcpy(s.val[0][0], full.data(), full.size() * sizeof(double));
    {
      std::ofstream out(outdir + "/ref.bck", std::ios::binary);
      s.save(out);
    }
    dump("bck_saved_interior", interior(s));
    std::cout << "BCKNAME " << name_file("press", ".bck", 12, boil::cart.iam()) << "\n";
    return 0;
  }
  if (mode == "bench-cg" || mode == "bench-ac" || mode == "bench-ic2") {
    if (argc < 11) { std::fprintf(stderr, "bench: too few arguments\n"); return 2; }
    int n = std::atoi(argv[2]);
    Case cs = parse_case(argv + 3, n, n, n);
    int lim = std::atoi(argv[10]);
    if (argc >= 14) {  // one sphere centred at (cx, cy, cz), radius rmin
      cs.centred = true;
      cs.cx = std::atof(argv[11]); cs.cy = std::atof(argv[12]); cs.cz = std::atof(argv[13]);
    }
    double t0 = now();
    World w = build_world(cs);
    double t1 = now();
    Scalar x(*w.d);
    neumann(x);
    x = 0.0;
    if (mode == "bench-cg" || mode == "bench-ic2") {
      CG ic2(*w.d, Prec::ic2());
      Krylov *sol = mode == "bench-cg" ? w.solver : &ic2;
      const char *tag = mode == "bench-cg" ? "cg" : "ic2";
      double ts = now();
      bool ok = sol->solve(w.pr->A, x, w.pr->fnew, MinIter(1), MaxIter(lim), tag,
                           ResRat(1e-8), ResTol(boil::femto), 1.0, -1, true);
      double te = now();
      std::cout << "BENCH " << tag << " n=" << n << " ok=" << ok << " setup_s=" << (t1 - t0)
                << " solve_s=" << (te - ts) << "\n";
    } else {
      AC ac(w.pr);
      ac.max_cycles(lim);
      w.pr->phi = 0.0;
      int ncyc = 0;
      double ts = now();
      bool ok = ac.vcycle(ResRat(1e-8), &ncyc);
      double te = now();
      std::cout << "BENCH ac n=" << n << " ok=" << ok << " cycles=" << ncyc
                << " setup_s=" << (t1 - t0) << " solve_s=" << (te - ts) << "\n";
    }
    return 0;
  }
  std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
  return 2;
}
#endif  // REF_DRIVER_EMBED
