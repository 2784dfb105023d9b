// config_tool.cpp -- host-only check of include/smac_config.hpp (no GPU):
//   config_tool <config.json> <outdir>
// parses and validates the configuration, prints its values as key=value
// lines, writes the initial alpha (cells) and v (y-faces) in the reference
// layout as raw float64 (alpha.f64, v.f64) and a VTI file result_000007.vti
// whose pressure is the cell index i + 1000 j + 1000000 k and whose alpha and
// velocity come from the initial fields. A configuration error prints the
// message on stderr and exits with status 2.
#include <cstdio>
#include <exception>
#include <vector>

#include "smac_config.hpp"

using namespace smac_config;

static void dump(const std::string& path, const std::vector<double>& a) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  std::fwrite(a.data(), sizeof(double), a.size(), f);
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: config_tool <config.json> <outdir>\n");
    return 1;
  }
  SimulationConfig c;
  try {
    c = load_config(argv[1]);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 2;
  }
  const std::string out = argv[2];
  const GridConfig& g = c.grid;
  std::printf("use_gpu=%d\n", c.hardware.use_gpu ? 1 : 0);
  std::printf("directory=%s\n", c.output.directory.c_str());
  std::printf("cells=%d,%d,%d\n", g.Nx, g.Ny, g.Nz);
  std::printf("size=%.17g,%.17g,%.17g\n", g.size_x, g.size_y, g.size_z);
  std::printf("origin=%.17g,%.17g,%.17g\n", g.origin_x_, g.origin_y_, g.origin_z_);
  std::printf("rho=%.17g,%.17g\n", c.fluid.phase0.density, c.fluid.phase1.density);
  std::printf("mu=%.17g,%.17g\n", c.fluid.phase0.viscosity, c.fluid.phase1.viscosity);
  std::printf("sigma=%.17g\n", c.fluid.surface_tension);
  std::printf("gravity=%.17g,%.17g,%.17g\n", c.fluid.gravity.x, c.fluid.gravity.y, c.fluid.gravity.z);
  std::printf("time=%.17g,%.17g,%.17g,%.17g,%.17g,%.17g\n", c.time.initial_dt, c.time.maximum_dt, c.time.end_time,
              c.time.interval, c.time.cfl, c.time.vof_cfl);
  std::printf("gmg_levels=%d\n", c.pressure_solver.gmg_levels);
  std::printf("tol=%.17g\n", c.pressure_solver.tol);
  std::printf("num_boundary_ids=%d\n", c.get_num_boundary_ids());
  std::printf("pure_neumann=%d\n", c.is_pure_neumann() ? 1 : 0);
  std::printf("alpha_substeps=%d\n", c.get_alpha_substeps());
  const BoundarySetup b = boundary_setup(c);
  std::printf("side_bcid=%d,%d,%d,%d,%d,%d\n", b.side_bcid[0], b.side_bcid[1], b.side_bcid[2], b.side_bcid[3],
              b.side_bcid[4], b.side_bcid[5]);
  for (const auto& e : b.table)
    std::printf("bc%d=%d,%.17g,%.17g,%.17g\n", e.id, (int)e.type, e.velocity.x, e.velocity.y, e.velocity.z);
  std::printf("initial_conditions=%zu\n", c.initial_conditions.size());
  const InitialFields f = initial_fields(c);
  dump(out + "/alpha.f64", f.alpha);
  dump(out + "/v.f64", f.v);
  // VTI: pressure = i + 1000 j + 1e6 k; u = w = 0 faces; v from the initial field
  const size_t sx = (size_t)g.Nx + 2, sy = (size_t)g.Ny + 2, sz = (size_t)g.Nz + 2;
  std::vector<double> p(sx * sy * sz, 0.0), u((sx + 1) * sy * sz, 0.0), w(sx * sy * (sz + 1), 0.0);
  for (size_t k = 0; k < sz; k++)
    for (size_t j = 0; j < sy; j++)
      for (size_t i = 0; i < sx; i++) p[(k * sy + j) * sx + i] = (double)i + 1000.0 * (double)j + 1e6 * (double)k;
  const std::vector<float> cells = pack_cells(g, p.data(), f.alpha.data(), u.data(), f.v.data(), w.data());
  write_vti(vti_filename(out, 7), g, cells.data());
  return 0;
}
