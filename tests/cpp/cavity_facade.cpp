// A driver shaped like the reference time loop (CFD/main.cu:225-317), built
// against include/smac_facade.hpp: lid-driven cavity, Jacobi PCG, the
// reference's GMG-PCG, or RBGS-multigrid PCG; pressure tolerance set per step
// through pressure_solver_->tol_.
// Usage: cavity_facade N STEPS OUTDIR [jacobi|gmg|rbgs]
// Writes u, v, w, p (reference layout, little-endian float64) to OUTDIR.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "smac_facade.hpp"

using namespace smac_facade;

static void dump(G_SMACSolver& g, int field, const std::string& path) {
  std::vector<double> h = g.download(field);
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  std::fwrite(h.data(), sizeof(double), h.size(), f);
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s N STEPS OUTDIR [jacobi|gmg|rbgs]\n", argv[0]);
    return 2;
  }
  const int n = std::atoi(argv[1]);
  const int steps = std::atoi(argv[2]);
  const std::string out = argv[3];
  const std::string mode = argc > 4 ? argv[4] : "jacobi";
  const bool use_gmg = mode == "gmg" || mode == "rbgs";
  try {
    G_SMACSolver g;
    g.set_calc_properties(0.0, 0.0, 0.0, 1.0, 1.0, 1.0, n, n, n);
    const int periodic[3] = {0, 0, 0};
    const unsigned char side[6] = {1, 1, 1, 2, 1, 1};  // lid on y-max
    g.set_boundaries(periodic, side,
                     {{SMAC_BC_NOSLIP, 0, 0, 0}, {SMAC_BC_NOSLIP, 0, 0, 0}, {SMAC_BC_NOSLIP, 1.0, 0, 0}});
    g.set_fluid(1.0, 1.0, 0.01, 0.01, 0.0, 0.0, 0.0);
    g.solver_malloc();

    G_PCGSolver pcg;
    G_GMGSolver gmg;
    gmg.initialize(g, 4);  // config.json "gmg_levels"
    gmg.rbgs_ = mode == "rbgs";
    pcg.set_gmg(gmg);
    pcg.set_solver(use_gmg ? PURENEUMANN_GMG_PCG : PURENEUMANN_STANDARD_PCG);
    g.pressure_solver_ = &pcg;

    SMACSolver solv;
    solv.dt_ = 1e-3;
    solv.inv_dt_ = 1.0 / solv.dt_;
    for (int step = 0; step < steps; step++) {
      g.pressure_solver_->tol_ = 1e-8;
      g.get_vof_vstar_rhouu_consistent(solv);
      g.update_vstar_boundary();
      g.solve_poisson();
      g.correct_vof_velocity(solv);
      g.update_boundary_ghost(solv);
      g.make_face_gradp();
      std::printf("step %d PCG iter = %d rel_res = %.3e\n", step, g.last_stats_.iters, g.last_stats_.rel_res);
    }
    dump(g, SMAC_FIELD_U, out + "/u.f64");
    dump(g, SMAC_FIELD_V, out + "/v.f64");
    dump(g, SMAC_FIELD_W, out + "/w.f64");
    dump(g, SMAC_FIELD_P, out + "/p.f64");
    dump(g, SMAC_FIELD_GRADPX, out + "/gradpx.f64");
    dump(g, SMAC_FIELD_GRADPY, out + "/gradpy.f64");
    dump(g, SMAC_FIELD_GRADPZ, out + "/gradpz.f64");
  } catch (const smac_error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
