// config_driver.cpp -- the reference's CFD time loop (CFD/main.cu:60-346) over
// the C++ facade, driven by a JSON configuration in the reference's schema:
//   config_driver <config.json> <outdir> <max_steps>
// Set-up from the configuration, CFL time step (CFDTime), VoF transport in
// alpha_substeps substeps, properties, predictor, reference-GMG PCG solve,
// correction, ghost update and face pressure gradient; result_%06d.vti at
// t = 0 and at every output step. After max_steps steps (or the end time) it
// writes the final fields as raw float64 (u, v, w, p, alpha) and the final VTI,
// and prints one line per step: step, dt, PCG iterations.
#include <cstdio>
#include <cstdlib>
#include <exception>
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
    std::fprintf(stderr, "usage: config_driver <config.json> <outdir> <max_steps>\n");
    return 1;
  }
  try {
    const smac_config::SimulationConfig config = smac_config::load_config(argv[1]);
    const std::string outdir = argv[2];
    const int max_steps = std::atoi(argv[3]);
    FileInOut fileIO(outdir);

    G_SMACSolver g_solv;
    setup_from_config(g_solv, config);
    SMACSolver solv;
    solv.gx_ = config.fluid.gravity.x;
    solv.gy_ = config.fluid.gravity.y;
    solv.gz_ = config.fluid.gravity.z;

    G_PCGSolver pcgSolver;
    g_solv.pressure_solver_ = &pcgSolver;
    G_GMGSolver gmgSolver;
    gmgSolver.initialize(g_solv, config.pressure_solver.gmg_levels);
    pcgSolver.set_gmg(gmgSolver);
    pcgSolver.set_solver(config.is_pure_neumann() ? PURENEUMANN_GMG_PCG : GMG_PCG);

    CFDTime cfdtime(config.time.initial_dt, config.time.maximum_dt, config.time.interval, config.time.end_time,
                    config.time.cfl, VARIBALE_TIME_STEP);
    const int alpha_substeps = config.get_alpha_substeps();
    fileIO.output_vti_binary_cellData(g_solv, 0);

    const double EPS = 1.E-12;
    int cur_step = 0;
    while (cfdtime.current_time_ < cfdtime.end_time_ - EPS && cur_step < max_steps) {
      g_solv.pressure_solver_->tol_ = cur_step == 0 ? 1e-11 : config.pressure_solver.tol;
      const double cfl = g_solv.calc_cfl();
      cfdtime.updateTime(cfl);
      solv.dt_ = cfdtime.dt_;
      solv.inv_dt_ = 1. / cfdtime.dt_;
      g_solv.dt_ = cfdtime.dt_;
      g_solv.inv_dt_ = 1. / cfdtime.dt_;

      g_solv.clear_alpha_flux_accum();
      const double sub_dt = cfdtime.dt_ / (double)alpha_substeps;
      for (int s = 0; s < alpha_substeps; s++) {
        g_solv.clear_alpha_flux();
        g_solv.alpha_flux_thincwlic_split(sub_dt, cur_step);
        g_solv.alpha_flux_accum();
      }
      g_solv.update_properties_by_alpha();
      g_solv.compute_mass_flux_from_alpha_flux(solv);
      g_solv.update_boundary_faces();
      g_solv.calc_surface_tension();

      g_solv.get_vof_vstar_rhouu_consistent(solv);
      g_solv.update_vstar_boundary();
      g_solv.solve_poisson();
      g_solv.correct_vof_velocity(solv);
      g_solv.update_boundary_ghost(solv);
      g_solv.make_face_gradp();
      std::printf("step %d dt %.17g iters %d\n", cur_step, cfdtime.dt_, g_solv.last_stats_.iters);
      if (cfdtime.isOutStep_) {
        fileIO.output_vti_binary_cellData(g_solv, cfdtime.current_steps_);
        cfdtime.isOutStep_ = false;
      }
      cur_step++;
    }
    dump(g_solv, SMAC_FIELD_U, outdir + "/u.f64");
    dump(g_solv, SMAC_FIELD_V, outdir + "/v.f64");
    dump(g_solv, SMAC_FIELD_W, outdir + "/w.f64");
    dump(g_solv, SMAC_FIELD_P, outdir + "/p.f64");
    dump(g_solv, SMAC_FIELD_ALPHA, outdir + "/alpha.f64");
    std::printf("final %s\n", fileIO.output_vti_binary_cellData(g_solv, 999999).c_str());
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 2;
  }
  return 0;
}
