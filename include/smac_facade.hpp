// smac_facade.hpp -- header-only C++ facade over the C-ABI (smac_c_api.h) that
// keeps the call surface of the reference's GPU solver classes, so a driver
// shaped like CFD/main.cu:225-317 compiles against it:
//
//   G_SMACSolver g;                     // CFD/G_SMACSolver.h:9-125
//   g.set_calc_properties(...);         // G_SMACSolver.cu:19-46
//   g.solver_malloc(bc, fluid);         // G_SMACSolver.cu:1451-1496 (+ host set-up of face/cell types)
//   G_PCGSolver pcg; G_GMGSolver gmg;   // pressure_solver/G_PCGSolver.h, G_GMGSolver.h
//   gmg.initialize(g, levels); pcg.set_gmg(gmg);
//   pcg.set_solver(PURENEUMANN_GMG_PCG);                  // G_PCGSolver.cu:592-613
//   g.pressure_solver_ = &pcg;
//   loop: g.pressure_solver_->tol_ = ...; g.calc_cfl(); g.get_vof_vstar_rhouu_consistent(solv);
//         g.update_vstar_boundary(); g.solve_poisson(); g.correct_vof_velocity(solv);
//
// Errors: the reference returns void and aborts; the facade throws smac_error
// carrying the C status code.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "smac_c_api.h"
#include "smac_config.hpp"

namespace smac_facade {

// PCG_TYPES of CFD/Enums.h:28-33
enum PCG_TYPES : unsigned char { STANDARD_PCG, PURENEUMANN_STANDARD_PCG, GMG_PCG, PURENEUMANN_GMG_PCG };

struct smac_error : std::runtime_error {
  int status;
  smac_error(int s, const std::string& what)
      : std::runtime_error(what + ": " + smac_strerror(s)), status(s) {}
};
inline void check(int s, const char* what) {
  if (s != SMAC_OK) throw smac_error(s, what);
}

// CFDTime (CFD/CFDTime.h, CFD/CFDTime.cpp:22-48): the CFL time-step rule of the
// reference loop. The first five steps take min(initdt, dt*cfl_thresh/(cfl+EPS));
// later steps take dt*cfl_thresh/(cfl+EPS), cut to reach the next output time
// and to max_dt; a step landing within 1e-8 of an output time marks it.
enum Time_mode { CONST_TIME_STEP, VARIBALE_TIME_STEP };
struct CFDTime {
  double dt_, initdt_, current_time_ = 0.0;
  int current_steps_ = 0;
  int outStepFreq_ = 0;
  double max_dt_, nextOut_, cfl_thresh_, end_time_, out_freq_time_;
  int mode_;
  bool isOutStep_ = false;
  CFDTime(double initdt, double max_dt, double outfreqtime, double end_time, double cfl_thresh, Time_mode mode)
      : dt_(initdt), initdt_(initdt), max_dt_(max_dt), nextOut_(outfreqtime), cfl_thresh_(cfl_thresh),
        end_time_(end_time), out_freq_time_(outfreqtime), mode_(mode) {}
  void updateTime(double cfl) {
    const double EPS = 1.E-12;  // CFD/hardCodedParameters.h:7
    if (current_steps_ < 5) {
      double dt_tmp = dt_ * cfl_thresh_ / (cfl + EPS);
      dt_ = initdt_ < dt_tmp ? initdt_ : dt_tmp;
    } else {
      dt_ = dt_ * cfl_thresh_ / (cfl + EPS);
      double till_out = nextOut_ - current_time_;
      if (dt_ >= till_out) dt_ = till_out;
      if (dt_ > max_dt_) dt_ = max_dt_;
      double t_next = current_time_ + dt_;
      if ((t_next - nextOut_ < 0 ? nextOut_ - t_next : t_next - nextOut_) <= 1e-8) {
        isOutStep_ = true;
        nextOut_ += out_freq_time_;
      }
    }
    current_time_ += dt_;
    current_steps_ += 1;
  }
};

// The host-side parameters the reference passes by value as SMACSolver (dt, gravity).
struct SMACSolver {
  double dt_ = 0.001;
  double inv_dt_ = 1000.0;
  double gx_ = 0.0, gy_ = 0.0, gz_ = 0.0;
};

struct G_SMACSolver;

// CFD/pressure_solver/G_PressureSolverBase.h:24-42
struct G_PressureSolverBase {
  virtual ~G_PressureSolverBase() {}
  virtual void solve(G_SMACSolver& solv) = 0;
  double tol_ = 1e-8;
};

// CFD/pressure_solver/G_GMGSolver.h:18-40. By default the V-cycle is the
// reference's own (weighted Jacobi, num_iter_fine_ / num_iter_coarse_ sweeps,
// same iterates); rbgs_ = true selects the red-black Gauss-Seidel V(nu1_,nu2_)
// multigrid instead (fewer PCG iterations, same solution to the tolerance).
struct G_GMGSolver {
  int num_levels_ = 4;  // requested levels (config "gmg_levels"); rbgs_: 0 = automatic
  int min_size_ = 6;    // fixed by the reference algorithm, informational
  int num_iter_coarse_ = 30;
  int num_iter_fine_ = 4;
  bool rbgs_ = false;
  int nu1_ = 2, nu2_ = 2;
  void initialize(G_SMACSolver&, int num_levels) { num_levels_ = num_levels; }
  void free_levels() {}
};

struct G_SMACSolver {
  smac_ctx* ctx_ = nullptr;
  smac_grid_desc grid_{};
  std::vector<smac_bc_entry> bc_;
  smac_fluid_desc fluid_{1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double dt_ = 0.001, inv_dt_ = 1000.0;
  double rho0_ = 1.0, rho1_ = 1.0, mu0_ = 0.0, mu1_ = 0.0;
  G_PressureSolverBase* pressure_solver_ = nullptr;  // non-owning, as the reference
  smac_solve_stats last_stats_{};

  G_SMACSolver() = default;
  G_SMACSolver(const G_SMACSolver&) = delete;
  G_SMACSolver& operator=(const G_SMACSolver&) = delete;
  ~G_SMACSolver() { solver_free(); }

  void set_calc_properties(double ox, double oy, double oz, double sx, double sy, double sz, int nx, int ny,
                           int nz) {
    grid_.nx = nx;
    grid_.ny = ny;
    grid_.nz = nz;
    grid_.sizex = sx;
    grid_.sizey = sy;
    grid_.sizez = sz;
    grid_.originx = ox;
    grid_.originy = oy;
    grid_.originz = oz;
  }
  // periodic directions (an extension) and the per-side bc ids / bc table
  // (StaggeredGrid::set_boundary_id, SimulationSetup::apply_boundary_conditions)
  void set_boundaries(const int periodic[3], const unsigned char side_bcid[6], const std::vector<smac_bc_entry>& bc) {
    for (int d = 0; d < 3; d++) grid_.periodic[d] = periodic[d];
    for (int s = 0; s < 6; s++) grid_.side_bcid[s] = side_bcid[s];
    bc_ = bc;
  }
  void set_fluid(double rho0, double rho1, double mu0, double mu1, double gx, double gy, double gz) {
    rho0_ = rho0;
    rho1_ = rho1;
    mu0_ = mu0;
    mu1_ = mu1;
    fluid_ = smac_fluid_desc{rho0, rho1, mu0, mu1, gx, gy, gz};
    if (ctx_) check(smac_set_fluid(ctx_, &fluid_), "smac_set_fluid");
  }
  void solver_malloc() {
    solver_free();
    grid_.num_bc = (int)bc_.size();
    grid_.bc_table = bc_.data();
    check(smac_create(&grid_, &fluid_, &ctx_), "smac_create");
  }
  void solver_free() {
    if (ctx_) smac_destroy(ctx_);
    ctx_ = nullptr;
  }
  // cpuTogpu / gpuTocpu for one field in the reference layout
  void upload(int field, const std::vector<double>& h) {
    check(smac_upload(ctx_, field, h.data(), h.size()), "smac_upload");
  }
  std::vector<double> download(int field) {
    std::vector<double> h(smac_field_count(ctx_, field));
    check(smac_download(ctx_, field, h.data(), h.size()), "smac_download");
    return h;
  }
  double calc_cfl() {
    double cc = 0, cv = 0;
    check(smac_calc_cfl(ctx_, dt_, &cc, &cv), "smac_calc_cfl");
    return cc > cv ? cc : cv;
  }
  void get_vof_vstar_rhouu_consistent(const SMACSolver& solv) {
    dt_ = solv.dt_;
    inv_dt_ = solv.inv_dt_;
    check(smac_predictor(ctx_, solv.dt_), "smac_predictor");
  }
  void update_vstar_boundary() {}  // fused into smac_predictor
  void solve_poisson() {
    if (!pressure_solver_) throw smac_error(SMAC_ERR_INVALID_ARG, "solve_poisson without pressure_solver_");
    pressure_solver_->solve(*this);
  }
  void correct_vof_velocity(const SMACSolver& solv) { check(smac_correct(ctx_, solv.dt_), "smac_correct"); }
  // VoF loop of CFD/main.cu:283-303: transport, properties, mass flux, wall faces
  void clear_alpha_flux_accum() { check(smac_clear_alpha_flux_accum(ctx_), "smac_clear_alpha_flux_accum"); }
  void clear_alpha_flux() {}  // each substep rewrites the face fluxes it computes
  void alpha_flux_thincwlic_split(double sub_dt, int step) {
    check(smac_alpha_flux_split(ctx_, sub_dt, step), "smac_alpha_flux_split");
  }
  void alpha_flux_accum() {}  // fused into the flux pass of smac_alpha_flux_split
  void update_properties_by_alpha() { check(smac_update_properties(ctx_, 0), "smac_update_properties"); }
  void compute_mass_flux_from_alpha_flux(const SMACSolver&) {}  // formed inside the predictor
  void update_boundary_faces() {}  // part of smac_update_properties
  void calc_surface_tension() {}   // surface tension is outside this path (sigma = 0)
  void update_boundary_ghost(const SMACSolver& solv) {
    check(smac_update_boundary_ghost(ctx_, solv.dt_), "smac_update_boundary_ghost");
  }
  void make_face_gradp() { check(smac_make_face_gradp(ctx_), "smac_make_face_gradp"); }
  // Two-way CFD-DEM loop (Coupling/main.cu:517-549): the coupling writes the
  // volume fractions and impulses (SMAC_FIELD_EPS, EPS_OLD, EPS_HALF, IMP{X,Y,Z})
  // with upload(); the mass flux is formed inside the two-way predictor.
  void compute_mass_flux_from_alpha_flux_two_way(const SMACSolver&) {}
  void get_vof_vstar_rhouu_consistent_two_way(const SMACSolver& solv) {
    dt_ = solv.dt_;
    inv_dt_ = solv.inv_dt_;
    check(smac_predictor_two_way(ctx_, solv.dt_), "smac_predictor_two_way");
  }
};

// Coupling/G_CFDDEMCoupling.h:149 -- the CFD side of the coupling object: the
// Poisson coefficients eps_f/rho_f of update_poisson_beta_two_way
// (Coupling/G_CFDDEMCoupling.cu:419-459). The DEM side (particles, drag,
// void-fraction mapping) is outside this path.
struct G_CFDDEMCoupling {
  void update_poisson_beta_two_way(G_SMACSolver& g) {
    check(smac_update_poisson_beta_two_way(g.ctx_), "smac_update_poisson_beta_two_way");
  }
};

// CFD/pressure_solver/G_PCGSolver.h:17-37. The stopping rule is the
// reference's: max|r|*dt < tol_, checked every 10 (Jacobi) or 5 (GMG)
// iterations, at most 2000 (G_PCGSolver.cu:318-327,612-625). The pure-Neumann
// treatment follows the boundary conditions, as SimulationConfig::is_pure_neumann;
// the PURENEUMANN_* and plain variants select the same solver here.
struct G_PCGSolver : G_PressureSolverBase {
  PCG_TYPES type_ = GMG_PCG;
  G_GMGSolver* gmg_ = nullptr;
  smac_solver_opts opts_{};
  G_PCGSolver() {
    smac_default_opts(&opts_);
    opts_.stop_rule = SMAC_STOP_REF_LINF_DT;
    opts_.max_iter = 2000;
  }
  void set_gmg(G_GMGSolver& gmg) { gmg_ = &gmg; }
  void set_solver(PCG_TYPES t) { type_ = t; }
  void solve(G_SMACSolver& s) override {
    smac_solver_opts o = opts_;
    o.tol = tol_;
    const bool gmg = type_ == GMG_PCG || type_ == PURENEUMANN_GMG_PCG;
    o.precond = SMAC_PRECOND_JACOBI;
    o.check_every = 10;
    if (gmg) {
      if (!gmg_) throw smac_error(SMAC_ERR_INVALID_ARG, "GMG_PCG without set_gmg");
      o.check_every = 5;
      o.precond = gmg_->rbgs_ ? SMAC_PRECOND_MG : SMAC_PRECOND_REF_GMG;
      o.mg_levels = gmg_->num_levels_;
      o.nu1 = gmg_->nu1_;
      o.nu2 = gmg_->nu2_;
      o.ref_sweeps = gmg_->num_iter_fine_;
      o.ref_coarse_sweeps = gmg_->num_iter_coarse_;
    }
    check(smac_pressure_solve(s.ctx_, s.dt_, &o, &s.last_stats_), "smac_pressure_solve");
  }
};

// The reference's set-up from its JSON configuration (CFD/main.cu:80-250 with
// SimulationSetup::apply_boundary_conditions / apply_initial_conditions):
// grid, boundaries, fluid, device arrays, the VoF initial fields, t = 0
// properties. Periodic directions are not expressible in the configuration.
inline void setup_from_config(G_SMACSolver& g, const smac_config::SimulationConfig& c) {
  const smac_config::GridConfig& gr = c.grid;
  g.set_calc_properties(gr.origin_x_, gr.origin_y_, gr.origin_z_, gr.size_x, gr.size_y, gr.size_z, gr.Nx, gr.Ny,
                        gr.Nz);
  const smac_config::BoundarySetup b = smac_config::boundary_setup(c);
  std::vector<smac_bc_entry> table;
  for (const auto& e : b.table) table.push_back(smac_bc_entry{(int32_t)e.type, e.velocity.x, e.velocity.y, e.velocity.z});
  const int periodic[3] = {0, 0, 0};
  g.set_boundaries(periodic, b.side_bcid, table);
  g.set_fluid(c.fluid.phase0.density, c.fluid.phase1.density, c.fluid.phase0.viscosity, c.fluid.phase1.viscosity,
              c.fluid.gravity.x, c.fluid.gravity.y, c.fluid.gravity.z);
  g.solver_malloc();
  const smac_config::InitialFields f = smac_config::initial_fields(c);
  g.upload(SMAC_FIELD_ALPHA, f.alpha);
  g.upload(SMAC_FIELD_V, f.v);
  check(smac_update_properties(g.ctx_, 1), "smac_update_properties");
}

// FileInOut (CFD/FileInOut.cpp:145-255): result_%06d.vti in outdir_, the cell
// data packed on the device (smac_cell_output)
struct FileInOut {
  std::string outdir_ = "results";
  explicit FileInOut(std::string outdir = "results") : outdir_(std::move(outdir)) {}
  std::string output_vti_binary_cellData(G_SMACSolver& g, int step) const {
    smac_config::GridConfig gr;
    gr.Nx = g.grid_.nx;
    gr.Ny = g.grid_.ny;
    gr.Nz = g.grid_.nz;
    gr.size_x = g.grid_.sizex;
    gr.size_y = g.grid_.sizey;
    gr.size_z = g.grid_.sizez;
    gr.origin_x_ = g.grid_.originx;
    gr.origin_y_ = g.grid_.originy;
    gr.origin_z_ = g.grid_.originz;
    const size_t n = (size_t)gr.Nx * (size_t)gr.Ny * (size_t)gr.Nz;
    std::vector<float> buf(5 * n);
    check(smac_cell_output(g.ctx_, buf.data(), buf.size()), "smac_cell_output");
    const std::string name = smac_config::vti_filename(outdir_, step);
    smac_config::write_vti(name, gr, buf.data());
    return name;
  }
};

}  // namespace smac_facade
