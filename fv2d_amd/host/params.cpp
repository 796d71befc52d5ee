// params.cpp -- readInifile / init_from_inifile / checkValidityIni restated (see params.h).
#include "params.h"

#include <cstring>
#include <map>
#include <ostream>

namespace fv2dh {

namespace {

// DeviceParams::init_from_inifile (SimInfo.h:356-459): same keys, defaults and order.
void read_device_params(RunParams &r) {
  Reader &rd = r.reader;
  r.Nx = rd.get_integer("mesh", "Nx", 32);
  r.Ny = rd.get_integer("mesh", "Ny", 32);
  r.Ng = rd.get_integer("mesh", "Nghosts", 2);
  r.xmin = rd.get_float("mesh", "xmin", 0.0);
  r.xmax = rd.get_float("mesh", "xmax", 1.0);
  r.ymin = rd.get_float("mesh", "ymin", 0.0);
  r.ymax = rd.get_float("mesh", "ymax", 1.0);
  r.Ntx = r.Nx + 2 * r.Ng;
  r.Nty = r.Ny + 2 * r.Ng;
  r.ibeg = r.Ng;
  r.iend = r.Ng + r.Nx;
  r.jbeg = r.Ng;
  r.jend = r.Ng + r.Ny;
  r.dx = (r.xmax - r.xmin) / r.Nx;
  r.dy = (r.ymax - r.ymin) / r.Ny;

  r.CFL = rd.get_float("solvers", "CFL", 0.8);
  const std::map<std::string, int> bc{
      {"reflecting", FV2D_BC_REFLECTING}, {"absorbing", FV2D_BC_ABSORBING}, {"periodic", FV2D_BC_PERIODIC}};
  r.boundary_x = rd.get_map_value(bc, "run", "boundaries_x", "reflecting");
  r.boundary_y = rd.get_map_value(bc, "run", "boundaries_y", "reflecting");
  const std::map<std::string, int> recon{{"pcm", FV2D_PCM}, {"pcm_wb", FV2D_PCM_WB}, {"plm", FV2D_PLM}};
  r.reconstruction = rd.get_map_value(recon, "solvers", "reconstruction", "pcm");
  const std::map<std::string, int> riemann{{"hll", FV2D_HLL}, {"hllc", FV2D_HLLC}, {"fslp", FV2D_FSLP}};
  r.riemann_solver = rd.get_map_value(riemann, "solvers", "riemann_solver", "hllc");

  r.epsilon = rd.get_float("misc", "epsilon", 1.0e-6);
  r.gamma0 = rd.get_float("physics", "gamma0", 5.0 / 3.0);
  r.m1 = rd.get_float("polytrope", "m1", 1.0);
  r.theta1 = rd.get_float("polytrope", "theta1", 10.0);
  r.m2 = rd.get_float("polytrope", "m2", 1.0);
  r.theta2 = rd.get_float("polytrope", "theta2", 10.0);
  r.well_balanced_flux_at_y_bc = rd.get_boolean("physics", "well_balanced_flux_at_y_bc", false);
  r.fslp_K = rd.get_float("physics", "fslp_K", 1.1);

  const std::map<std::string, int> grav{
      {"none", FV2D_GRAV_NONE}, {"constant", FV2D_GRAV_CONSTANT}, {"analytical", FV2D_GRAV_ANALYTICAL}};
  r.gravity_mode = rd.get_map_value(grav, "gravity", "mode", "none");
  r.gx = rd.get_float("gravity", "gx", 0.0);
  r.gy = rd.get_float("gravity", "gy", 0.0);
  const std::map<std::string, int> agm{{"hot_bubble", 0}};
  r.analytical_gravity_mode = rd.get_map_value(agm, "gravity", "analytical_mode", "hot_bubble");

  r.thermal_conductivity_active = rd.get_boolean("thermal_conduction", "active", false);
  const std::map<std::string, int> tcm{{"constant", FV2D_TCM_CONSTANT}, {"B02", FV2D_TCM_B02}};
  r.thermal_conductivity_mode = rd.get_map_value(tcm, "thermal_conduction", "conductivity_mode", "constant");
  r.kappa = rd.get_float("thermal_conduction", "kappa", 0.0);
  const std::map<std::string, int> bctc{{"none", FV2D_BCTC_NONE},
                                        {"fixed_temperature", FV2D_BCTC_FIXED_TEMPERATURE},
                                        {"fixed_gradient", FV2D_BCTC_FIXED_GRADIENT}};
  r.bctc_ymin = rd.get_map_value(bctc, "thermal_conduction", "bc_ymin", "none");
  r.bctc_ymax = rd.get_map_value(bctc, "thermal_conduction", "bc_ymax", "none");
  r.bctc_ymin_value = rd.get_float("thermal_conduction", "bc_ymin_value", 1.0);
  r.bctc_ymax_value = rd.get_float("thermal_conduction", "bc_ymax_value", 1.0);

  r.viscosity_active = rd.get_boolean("viscosity", "active", false);
  const std::map<std::string, int> vm{{"constant", 0}};
  r.viscosity_mode = rd.get_map_value(vm, "viscosity", "viscosity_mode", "constant");
  r.mu = rd.get_float("viscosity", "mu", 0.0);

  r.h84_pert = rd.get_float("H84", "perturbation", 1.0e-4);
  r.c91_pert = rd.get_float("C91", "perturbation", 1.0e-3);
  r.hot_bubble_g0 = rd.get_float("hot_bubble", "g0", 0.0);

  r.kh_a = rd.get_float("kelvin_helmholtz", "a", 0.05);
  r.kh_amp = rd.get_float("kelvin_helmholtz", "amp", 0.01);
  r.kh_P0 = rd.get_float("kelvin_helmholtz", "P0", 1.0);
  r.kh_rho_fac = rd.get_float("kelvin_helmholtz", "rho_fac", 0.0);
  r.kh_sigma = rd.get_float("kelvin_helmholtz", "sigma", 0.2);
  // the reference reads these three from a section spelled "kelvin_helmholts"
  r.kh_uflow = rd.get_float("kelvin_helmholts", "uflow", 1.0);
  r.kh_y1 = rd.get_float("kelvin_helmholts", "y1", 0.5);
  r.kh_y2 = rd.get_float("kelvin_helmholts", "y2", 1.5);

  r.gresho_density = rd.get_float("gresho_vortex", "density", 1.0);
  r.gresho_Mach = rd.get_float("gresho_vortex", "Mach", 0.1);
}

RunParams read_all(Reader reader, std::ostream &warn) {
  RunParams r;
  r.reader = std::move(reader);
  Reader &rd = r.reader;
  r.tend = rd.get_float("run", "tend", 1.0);
  r.multiple_outputs = rd.get_boolean("run", "multiple_outputs", false);
  r.restart_file = rd.get("run", "restart_file", "");
  r.save_freq = rd.get_float("run", "save_freq", 1.0e-1);
  r.filename_out = rd.get("run", "output_filename", "run");
  r.output_path = rd.get("run", "output_path", "./");
  const std::map<std::string, int> ts{{"euler", FV2D_TS_EULER}, {"RK2", FV2D_TS_RK2}};
  r.time_stepping = rd.get_map_value(ts, "solvers", "time_stepping", "euler");
  r.problem = rd.get("physics", "problem", "blast");
  r.seed = rd.get_integer("misc", "seed", 12345);
  r.log_frequency = rd.get_integer("misc", "log_frequency", 10);
  r.epsilon_reset_negative = rd.get_float("misc", "epsilon_reset_negative", 1.0e-8);
  read_device_params(r);
  check_validity(r, warn);
  return r;
}

}  // namespace

void check_validity(const RunParams &p, std::ostream &warn) {
  const auto &sections = p.reader.ini.sections();  // as written in the file
  const auto &kv = p.reader.ini.values();           // "section=key" (lower case)
  const auto &known = p.reader.registered;          // lower case
  for (const std::string &s : sections) {
    auto ks = known.find(s);
    if (ks == known.end()) {
      warn << "WARNING: section [" << s << "] is unknown." << std::endl;
      continue;
    }
    const std::string prefix = s + "=";
    for (const auto &e : kv) {
      if (e.first.compare(0, prefix.size(), prefix) == 0) {
        const std::string key = e.first.substr(prefix.size());
        if (!ks->second.count(key))
          warn << "WARNING: parameter `" << key << "` in section [" << s << "] is unknown." << std::endl;
      }
    }
  }
}

RunParams read_ini(const std::string &filename, std::ostream &warn) { return read_all(Reader(filename), warn); }

RunParams read_ini_text(const std::string &text, std::ostream &warn) {
  return read_all(Reader(IniFile::from_string(text)), warn);
}

InitType init_type(const std::string &problem) {
  static const std::map<std::string, InitType> m{{"sod_x", SOD_X},
                                                 {"sod_y", SOD_Y},
                                                 {"blast", BLAST},
                                                 {"rayleigh-taylor", RAYLEIGH_TAYLOR},
                                                 {"diffusion", DIFFUSION},
                                                 {"H84", H84},
                                                 {"C91", C91},
                                                 {"kelvin_helmholtz", KELVIN_HELMHOLTZ},
                                                 {"gresho_vortex", GRESHO_VORTEX}};
  auto it = m.find(problem);
  if (it == m.end()) throw std::runtime_error("Error unknown problem " + problem);
  return it->second;
}

fv2d_params to_fv2d_params(const RunParams &r, int precision) {
  fv2d_params p;
  std::memset(&p, 0, sizeof p);
  p.Nx = r.Nx;
  p.Ny = r.Ny;
  p.Ng = r.Ng;
  p.xmin = r.xmin;
  p.xmax = r.xmax;
  p.ymin = r.ymin;
  p.ymax = r.ymax;
  p.CFL = r.CFL;
  p.reconstruction = r.reconstruction;
  p.riemann_solver = r.riemann_solver;
  p.time_stepping = r.time_stepping;
  p.boundary_x = r.boundary_x;
  p.boundary_y = r.boundary_y;
  p.gamma0 = r.gamma0;
  p.well_balanced_flux_at_y_bc = r.well_balanced_flux_at_y_bc ? 1 : 0;
  p.fslp_K = r.fslp_K;
  p.gravity_mode = r.gravity_mode;
  p.gx = r.gx;
  p.gy = r.gy;
  p.hot_bubble_g0 = r.hot_bubble_g0;
  p.tc_active = r.thermal_conductivity_active ? 1 : 0;
  p.tc_mode = r.thermal_conductivity_mode;
  p.kappa = r.kappa;
  p.bctc_ymin = r.bctc_ymin;
  p.bctc_ymax = r.bctc_ymax;
  p.bctc_ymin_value = r.bctc_ymin_value;
  p.bctc_ymax_value = r.bctc_ymax_value;
  p.visc_active = r.viscosity_active ? 1 : 0;
  p.mu = r.mu;
  p.epsilon = r.epsilon;
  p.epsilon_reset_negative = r.epsilon_reset_negative;
  p.precision = precision;
  p.seg_rows = 0;
  return p;
}

}  // namespace fv2dh
