// psb_gen.cpp -- synthetic pressure systems for the benchmark cases.
//
// Builds the variable-density pressure matrix and right-hand side the way
// PSI-BOIL's Pressure does on a Cartesian grid with Neumann walls:
//   face coefficient a = dS / (d * rho_face), rho_face = arithmetic mean of
//   the two cell densities   (pressure_discretize.cpp:24-57,
//                             Src/Matter/property_value.cpp:18-32)
//   c = sum of the six couplings, then for each wall face c -= a, a = 0
//                            (pressure_discretize.cpp:65-67,
//                             centered_system_bnd.cpp:27-61)
//   solid ("off") cells decoupled on both sides, c *= 1e-12 (:93-148)
//   b = divergence of the face velocities times dS (pressure_update_rhs.cpp
//   :27-52), 0 in solid cells, shifted to zero mean over the fluid cells.
// Ghost cells follow Grid1D's default BndGrid::wall() rule
// (Src/Grid/grid1d_correct_boundaries.cpp): the first and last cell-centre
// distances run to the wall.  Random numbers are splitmix64, u01 = (x>>11)
// * 2^-53, so every consumer of a seed sees the same case.
//
// Host code; the large loops run on std::threads.  Every per-cell
// expression keeps one fixed evaluation order (built without FMA
// contraction), so results do not depend on the thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/psb.h"

namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// n-th draw (0-based) of the stream seeded with `seed`
inline double u01_at(uint64_t seed, uint64_t n) {
  return (double)(mix64(seed + (n + 1) * kGamma) >> 11) * (1.0 / 9007199254740992.0);
}

template <typename F>
void parallel_for(long long n, F f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < (1 << 16)) nt = 1;
  std::vector<std::thread> th;
  long long per = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; t++) {
    long long a = t * per, b = std::min(n, a + per);
    if (a >= b) break;
    th.emplace_back([=]() {
      for (long long q = a; q < b; q++) f(q);
    });
  }
  for (auto &t : th) t.join();
}

struct Axis {
  int n;
  std::vector<double> xn, xc, dxc, dxw, dxe;
  void metrics() {
    xc.resize(n); dxc.resize(n); dxw.resize(n); dxe.resize(n);
    for (int i = 0; i < n; i++) {
      xc[i] = 0.5 * (xn[i + 1] + xn[i]);
      dxc[i] = xn[i + 1] - xn[i];
    }
    for (int i = 0; i < n; i++) {
      double xm = (i == 0) ? 0.5 * (xn[0] + xn[0]) : xc[i - 1];
      double xp = (i == n - 1) ? 0.5 * (xn[n] + xn[n]) : xc[i + 1];
      dxw[i] = xc[i] - xm;
      dxe[i] = xp - xc[i];
    }
  }
};

Axis uniform_axis(int n, double h) {
  Axis a;
  a.n = n;
  a.xn.resize(n + 1);
  for (int i = 0; i <= n; i++) a.xn[i] = h * i;
  a.metrics();
  return a;
}

// tanh clustering towards both walls: z_k = L/2 (1 + tanh(beta (2k/n - 1)) / tanh beta)
Axis tanh_axis(int n, double L, double beta) {
  Axis a;
  a.n = n;
  a.xn.resize(n + 1);
  const double tb = std::tanh(beta);
  for (int k = 0; k <= n; k++)
    a.xn[k] = 0.5 * L * (1.0 + std::tanh(beta * (2.0 * k / n - 1.0)) / tb);
  a.xn[0] = 0.0;
  a.xn[n] = L;
  a.metrics();
  return a;
}

}  // namespace

extern "C" int psb_generate(const psb_case *cs, double *c, double *w, double *e, double *s,
                            double *n, double *b_, double *t, double *rhs, unsigned char *off_out,
                            double *dxc_out, double *dyc_out, double *dzc_out) {
  if (!cs || !c || !w || !e || !s || !n || !b_ || !t || !rhs) return PSB_EINVAL;
  const int nx = cs->nx, ny = cs->ny, nz = cs->nz;
  if (nx < 1 || ny < 1 || nz < 1 || cs->nsph < 0) return PSB_EINVAL;
  if ((long long)nx * ny * nz >= (1LL << 31)) return PSB_EINVAL;
  const int nmax = std::max(nx, std::max(ny, nz));
  const double h = 1.0 / nmax;
  Axis X = uniform_axis(nx, h), Y = uniform_axis(ny, h);
  Axis Z = cs->stretch_z > 0.0 ? tanh_axis(nz, cs->lz > 0 ? cs->lz : nz * h, cs->stretch_z)
                               : uniform_axis(nz, h);
  const double Lx = X.xn[nx], Ly = Y.xn[ny], Lz = Z.xn[nz];
  const long long N = (long long)nx * ny * nz;
  const int gy = ny + 2, gz = nz + 2;
  auto id = [&](int i, int j, int k) { return ((long long)i * ny + j) * nz + k; };

  // ---- spheres: draw order r, cx, cy, cz per sphere -----------------------
  std::vector<double> sph;
  if (cs->centred) {
    sph = {cs->cx, cs->cy, cs->cz, cs->rmin};
  } else {
    uint64_t q = 0;
    for (int p = 0; p < cs->nsph; p++) {
      double r = cs->rmin + (cs->rmax - cs->rmin) * u01_at(cs->seed_sph, q++);
      double cx = r + (Lx - 2.0 * r) * u01_at(cs->seed_sph, q++);
      double cy = r + (Ly - 2.0 * r) * u01_at(cs->seed_sph, q++);
      double cz = r + (Lz - 2.0 * r) * u01_at(cs->seed_sph, q++);
      sph.insert(sph.end(), {cx, cy, cz, r});
    }
  }
  // ---- cell density with a one-cell ghost layer (Neumann colour) ---------
  std::vector<uint8_t> inside((size_t)N, 0);
  const int nsp = (int)sph.size() / 4;
  for (int p = 0; p < nsp; p++) {
    const double cx = sph[4 * p], cy = sph[4 * p + 1], cz = sph[4 * p + 2], r = sph[4 * p + 3];
    // candidate index range: every cell centre within r lies inside it
    auto range = [&](const Axis &A, double c0) {
      int lo = (int)(std::lower_bound(A.xc.begin(), A.xc.end(), c0 - r) - A.xc.begin()) - 1;
      int hi = (int)(std::upper_bound(A.xc.begin(), A.xc.end(), c0 + r) - A.xc.begin()) + 1;
      return std::make_pair(std::max(lo, 0), std::min(hi, A.n));
    };
    auto ri = range(X, cx), rj = range(Y, cy), rk = range(Z, cz);
    for (int i = ri.first; i < ri.second; i++)
      for (int j = rj.first; j < rj.second; j++)
        for (int k = rk.first; k < rk.second; k++) {
          double dx = X.xc[i] - cx, dy = Y.xc[j] - cy, dz = Z.xc[k] - cz;
          if (std::sqrt(dx * dx + dy * dy + dz * dz) < r) inside[id(i, j, k)] = 1;
        }
  }
  std::vector<double> rho((size_t)(nx + 2) * gy * gz);
  parallel_for(nx + 2, [&](long long ii) {
    const int i = (int)ii - 1;
    const int ci = std::min(std::max(i, 0), nx - 1);
    for (int j = -1; j <= ny; j++) {
      const int cj = std::min(std::max(j, 0), ny - 1);
      for (int k = -1; k <= nz; k++) {
        const int ck = std::min(std::max(k, 0), nz - 1);
        rho[((size_t)(i + 1) * gy + (j + 1)) * gz + (k + 1)] =
            inside[id(ci, cj, ck)] ? cs->rho_in : cs->rho_out;
      }
    }
  });
  // ---- solid cylinder along y (off cells) --------------------------------
  std::vector<uint8_t> off;
  if (cs->cylinder) {
    off.assign((size_t)N, 0);
    const double R = 0.5 * cs->cyl_d * Lx, cx = 0.5 * Lx, cz = 0.5 * Lz;
    parallel_for(nx, [&](long long i) {
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          double dx = X.xc[i] - cx, dz = Z.xc[k] - cz;
          if (std::sqrt(dx * dx + dz * dz) < R) off[id((int)i, j, k)] = 1;
        }
    });
  }
#define RHO(i, j, k) rho[((size_t)((i) + 1) * gy + ((j) + 1)) * gz + ((k) + 1)]
  std::vector<double> csum;
  if (!off.empty()) csum.resize((size_t)N);
  // neighbouring coefficients and their sum
  parallel_for(nx, [&](long long ii) {
    const int i = (int)ii;
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        const long long P = id(i, j, k);
        const double rp = RHO(i, j, k);
        const double a_x = Y.dxc[j] * Z.dxc[k];
        const double a_y = X.dxc[i] * Z.dxc[k];
        const double a_z = X.dxc[i] * Y.dxc[j];
        w[P] = a_x / X.dxw[i] / (0.5 * (rp + RHO(i - 1, j, k)));
        e[P] = a_x / X.dxe[i] / (0.5 * (RHO(i + 1, j, k) + rp));
        s[P] = a_y / Y.dxw[j] / (0.5 * (rp + RHO(i, j - 1, k)));
        n[P] = a_y / Y.dxe[j] / (0.5 * (RHO(i, j + 1, k) + rp));
        b_[P] = a_z / Z.dxw[k] / (0.5 * (rp + RHO(i, j, k - 1)));
        t[P] = a_z / Z.dxe[k] / (0.5 * (RHO(i, j, k + 1) + rp));
        c[P] = w[P] + e[P] + s[P] + n[P] + b_[P] + t[P];
        if (!csum.empty()) csum[P] = c[P];
      }
  });
#undef RHO
  if (!off.empty()) {
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          const long long P = id(i, j, k);
          if (!off[P]) continue;
          w[P] = e[P] = s[P] = n[P] = b_[P] = t[P] = 0.0;
          if (i > 0) e[id(i - 1, j, k)] = 0.0;
          if (i < nx - 1) w[id(i + 1, j, k)] = 0.0;
          if (j > 0) n[id(i, j - 1, k)] = 0.0;
          if (j < ny - 1) s[id(i, j + 1, k)] = 0.0;
          if (k > 0) t[id(i, j, k - 1)] = 0.0;
          if (k < nz - 1) b_[id(i, j, k + 1)] = 0.0;
        }
    parallel_for(N, [&](long long P) {
      c[P] = off[P] ? csum[P] : w[P] + e[P] + s[P] + n[P] + b_[P] + t[P];
    });
  }
  // Neumann walls: imin, imax, jmin, jmax, kmin, kmax
  for (int j = 0; j < ny; j++)
    for (int k = 0; k < nz; k++) {
      long long P = id(0, j, k);
      c[P] -= w[P]; w[P] = 0.0;
    }
  for (int j = 0; j < ny; j++)
    for (int k = 0; k < nz; k++) {
      long long P = id(nx - 1, j, k);
      c[P] -= e[P]; e[P] = 0.0;
    }
  for (int i = 0; i < nx; i++)
    for (int k = 0; k < nz; k++) {
      long long P = id(i, 0, k);
      c[P] -= s[P]; s[P] = 0.0;
    }
  for (int i = 0; i < nx; i++)
    for (int k = 0; k < nz; k++) {
      long long P = id(i, ny - 1, k);
      c[P] -= n[P]; n[P] = 0.0;
    }
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      long long P = id(i, j, 0);
      c[P] -= b_[P]; b_[P] = 0.0;
    }
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      long long P = id(i, j, nz - 1);
      c[P] -= t[P]; t[P] = 0.0;
    }
  if (!off.empty())
    parallel_for(N, [&](long long P) {
      if (off[P]) c[P] *= 1.0e-12;
    });

  // ---- right-hand side -----------------------------------------------------
  // face velocities: u faces, then v, then w, each in storage order, uniform
  // (-1,1); wall-normal faces 0 (their draw is still consumed)
  const uint64_t nu = (uint64_t)(nx + 1) * ny * nz, nv = (uint64_t)nx * (ny + 1) * nz;
  const uint64_t seed = cs->seed_vel;
  auto uf = [&](int i, int j, int k) -> double {
    if (i == 0 || i == nx) return 0.0;
    return 2.0 * u01_at(seed, ((uint64_t)i * ny + j) * nz + k) - 1.0;
  };
  auto vf = [&](int i, int j, int k) -> double {
    if (j == 0 || j == ny) return 0.0;
    return 2.0 * u01_at(seed, nu + ((uint64_t)i * (ny + 1) + j) * nz + k) - 1.0;
  };
  auto wf = [&](int i, int j, int k) -> double {
    if (k == 0 || k == nz) return 0.0;
    return 2.0 * u01_at(seed, nu + nv + ((uint64_t)i * ny + j) * (nz + 1) + k) - 1.0;
  };
  parallel_for(nx, [&](long long ii) {
    const int i = (int)ii;
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        const long long P = id(i, j, k);
        const double a_x = Y.dxc[j] * Z.dxc[k] * 1.0;
        const double a_y = X.dxc[i] * Z.dxc[k] * 1.0;
        const double a_z = X.dxc[i] * Y.dxc[j] * 1.0;
        double f = 0.0;
        f += a_x * uf(i, j, k);
        f -= a_x * uf(i + 1, j, k);
        f += a_y * vf(i, j, k);
        f -= a_y * vf(i, j + 1, k);
        f += a_z * wf(i, j, k);
        f -= a_z * wf(i, j, k + 1);
        if (!off.empty() && off[P]) f = 0.0;
        rhs[P] = f;
      }
  });
  // zero mean over fluid cells, summed in cell order
  double sum = 0.0;
  long long cnt = 0;
  for (long long P = 0; P < N; P++)
    if (off.empty() || !off[P]) {
      sum += rhs[P];
      cnt++;
    }
  if (cnt > 0) {
    const double mean = sum / (double)cnt;
    parallel_for(N, [&](long long P) {
      if (off.empty() || !off[P]) rhs[P] -= mean;
    });
  }
  if (off_out) {
    if (off.empty())
      std::memset(off_out, 0, (size_t)N);
    else
      std::memcpy(off_out, off.data(), (size_t)N);
  }
  if (dxc_out) std::memcpy(dxc_out, X.dxc.data(), nx * sizeof(double));
  if (dyc_out) std::memcpy(dyc_out, Y.dxc.data(), ny * sizeof(double));
  if (dzc_out) std::memcpy(dzc_out, Z.dxc.data(), nz * sizeof(double));
  return PSB_OK;
}

// Batch of pressure cases: per system, each dimension uniform in
// {32, 48, 64, 96, 128}, 1-8 bubbles with radius uniform in 0.05-0.2 of the
// shortest side, density ratio uniform in 10-1000 (bubbles of density 1 in
// liquid of density `ratio`), and its own seeds for the bubbles and the
// face velocities.  Draws come from one splitmix64 stream of `seed`.
extern "C" int psb_batch_cases(unsigned long long seed, int nsys, psb_case *out) {
  if (nsys < 0 || (nsys > 0 && !out)) return PSB_EINVAL;
  static const int dims[5] = {32, 48, 64, 96, 128};
  uint64_t q = 0;
  for (int s = 0; s < nsys; s++) {
    psb_case &c = out[s];
    std::memset(&c, 0, sizeof(c));
    c.nx = dims[std::min(4, (int)(5.0 * u01_at(seed, q++)))];
    c.ny = dims[std::min(4, (int)(5.0 * u01_at(seed, q++)))];
    c.nz = dims[std::min(4, (int)(5.0 * u01_at(seed, q++)))];
    c.nsph = 1 + std::min(7, (int)(8.0 * u01_at(seed, q++)));
    const int nmax = std::max(c.nx, std::max(c.ny, c.nz));
    const int nmin = std::min(c.nx, std::min(c.ny, c.nz));
    const double side = (double)nmin / nmax;  // h = 1 / max(n)
    c.rmin = 0.05 * side;
    c.rmax = 0.2 * side;
    c.rho_in = 1.0;
    c.rho_out = 10.0 + 990.0 * u01_at(seed, q++);
    c.seed_sph = mix64(seed + 0x1000 + 2 * (uint64_t)s);
    c.seed_vel = mix64(seed + 0x1001 + 2 * (uint64_t)s);
  }
  return PSB_OK;
}
