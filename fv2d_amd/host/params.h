// params.h -- run parameters read from a .ini file: the reference's DeviceParams and
// Params (SimInfo.h:266-492), readInifile (SimInfo.h:529-568) and checkValidityIni
// (SimInfo.h:501-527), and their translation to the C ABI's fv2d_params.
//
// Every parameter is read with the reference's section, key, default and reader type, in
// its order, so that the set of known keys (unknown-key warnings), the last.ini dump and
// the values are the same. Floating-point parameters go through single precision
// (INIReader::GetFloat), as in the reference: CFL = 0.8 in a file is 0.800000011920929.
#pragma once

#include <string>

#include "../../include/fv2d_amd.h"
#include "ini.h"

namespace fv2dh {

enum InitType { SOD_X, SOD_Y, BLAST, RAYLEIGH_TAYLOR, DIFFUSION, H84, C91, KELVIN_HELMHOLTZ, GRESHO_VORTEX };

struct RunParams {
  // [run]
  double tend = 1.0, save_freq = 0.1;
  bool multiple_outputs = false;
  std::string restart_file, filename_out = "run", output_path = "./";
  int time_stepping = FV2D_TS_EULER;
  std::string problem = "blast";
  // [misc]
  int seed = 12345, log_frequency = 10;
  double epsilon_reset_negative = 1e-8;
  double epsilon = 1e-6;
  // [mesh]
  int Nx = 32, Ny = 32, Ng = 2;
  double xmin = 0, xmax = 1, ymin = 0, ymax = 1;
  int Ntx = 0, Nty = 0, ibeg = 0, iend = 0, jbeg = 0, jend = 0;
  double dx = 0, dy = 0;
  // [solvers]
  double CFL = 0.8;
  int boundary_x = FV2D_BC_REFLECTING, boundary_y = FV2D_BC_REFLECTING;
  int reconstruction = FV2D_PCM, riemann_solver = FV2D_HLLC;
  // [physics] / [polytrope]
  double gamma0 = 5.0 / 3.0, m1 = 1, theta1 = 10, m2 = 1, theta2 = 10;
  bool well_balanced_flux_at_y_bc = false;
  double fslp_K = 1.1;
  // [gravity]
  int gravity_mode = FV2D_GRAV_NONE, analytical_gravity_mode = 0;
  double gx = 0, gy = 0;
  // [thermal_conduction]
  bool thermal_conductivity_active = false;
  int thermal_conductivity_mode = FV2D_TCM_CONSTANT;
  double kappa = 0;
  int bctc_ymin = FV2D_BCTC_NONE, bctc_ymax = FV2D_BCTC_NONE;
  double bctc_ymin_value = 1, bctc_ymax_value = 1;
  // [viscosity]
  bool viscosity_active = false;
  int viscosity_mode = 0;
  double mu = 0;
  // problem parameters
  double h84_pert = 1e-4, c91_pert = 1e-3, hot_bubble_g0 = 0;
  double kh_a = 0.05, kh_amp = 0.01, kh_P0 = 1, kh_rho_fac = 0, kh_sigma = 0.2, kh_uflow = 1, kh_y1 = 0.5, kh_y2 = 1.5;
  double gresho_density = 1, gresho_Mach = 0.1;

  Reader reader;
};

// Reads every parameter (readInifile); throws std::runtime_error on a bad enum value.
// Unknown sections and keys are reported on `warn` (checkValidityIni).
RunParams read_ini(const std::string &filename, std::ostream &warn);
RunParams read_ini_text(const std::string &text, std::ostream &warn);

// The unknown-section / unknown-key warnings of checkValidityIni.
void check_validity(const RunParams &p, std::ostream &warn);

// InitFunctor's problem table (Init.h:292-305); throws on an unknown problem.
InitType init_type(const std::string &problem);

// The C ABI parameters of this run.
fv2d_params to_fv2d_params(const RunParams &p, int precision);

}  // namespace fv2dh
