// gk_driver.cpp -- runs the reference's own callers of Linear::solve with the
// GpuKrylov binding (include/psb_krylov.hpp) in place of its CG, beside the
// same callers with the reference's CG on the CPU, on one benchmark case.
//
// TEST INFRASTRUCTURE ONLY (built by oracle/ref.mk against the reference's
// objects and libpsb.so; tests/test_gpu_binding.py runs it on the GPU).
//
//   gk_driver centered OUT n nsph rmin rmax rho_in rho_out seed_sph seed_vel
//       Centered::solve (centered_solve.cpp:9-40: update_rhs, then the
//       solver with scale * dti) of the case's Pressure to ResRat 1e-8
//   gk_driver ac OUT n nsph rmin rmax rho_in rho_out seed_sph seed_vel
//       AC::vcycle (additive.h:39-41) to ResRat 1e-6: every level's
//       smoother and the coarsest solve call the equation's solver on that
//       level's matrix (additive_cycle_functions.cpp:89-107)
//
// Each mode writes <tag>_phi.f64 for tag cpu and gpu, and prints the solver
// log lines and "CONTEXTS n" (the library contexts GpuKrylov holds).
#define REF_DRIVER_EMBED
#include "ref_driver.cpp"

#include "psb_krylov.hpp"

namespace {

Case parse(char **argv) {
  const int n = std::atoi(argv[3]);
  return parse_case(argv + 4, n, n, n);
}

Krylov *make_gpu(const Domain &d) { return new GpuKrylov(d); }

}  // namespace

int main(int argc, char **argv) {
  std::cout << std::setprecision(17);
  if (argc < 11) {
    std::fprintf(stderr, "usage: gk_driver centered|ac OUT n nsph rmin rmax rho_in rho_out "
                         "seed_sph seed_vel\n");
    return 2;
  }
  const std::string mode = argv[1];
  outdir = argv[2];
  const Case cs = parse(argv);
  for (int gpu = 0; gpu < 2; gpu++) {
    const char *tag = gpu ? "gpu" : "cpu";
    // every level of the Pressure (and so AC) holds this one solver object
    World w = build_world(cs, gpu ? make_gpu : nullptr);
    GpuKrylov *gk = gpu ? static_cast<GpuKrylov *>(w.solver) : nullptr;
    w.pr->phi = 0.0;
    std::cout << "RUN " << tag << std::endl;
    if (mode == "centered") {
      w.pr->solve(ResTol(boil::femto), ResRat(1e-8), "centered", 2000);
    } else if (mode == "ac") {
      AC ac(w.pr);
      ac.max_cycles(20);
      int ncyc = 0;
      const bool ok = ac.vcycle(ResRat(1e-6), &ncyc);
      std::cout << "ACRET " << tag << " " << ok << " " << ncyc << std::endl;
    } else {
      std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
      return 2;
    }
    dump(std::string(tag) + "_phi", interior(w.pr->phi));
    if (gk) {
      std::cout << "CONTEXTS " << gk->contexts() << " STATUS " << gk->last_status()
                << std::endl;
    }
  }
  return 0;
}
