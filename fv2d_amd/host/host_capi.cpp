// host_capi.cpp -- C entry points of libfv2d_host.so for the tests and Python tools: the
// parameters, warnings and last.ini text of a .ini file, the initial state of its problem,
// snapshot writing / restart loading without a GPU, and HDF5 read-back (h5py is not part
// of this environment).
#include <cstring>
#include <iomanip>
#include <sstream>
#include <string>

#include "h5.h"
#include "init.h"
#include "io.h"
#include "params.h"

using namespace fv2dh;

namespace {

thread_local std::string g_err;

int copy_out(const std::string &s, char *out, int cap) {
  if (out && cap > 0) {
    const size_t n = std::min(s.size(), (size_t)cap - 1);
    std::memcpy(out, s.data(), n);
    out[n] = 0;
  }
  return (int)s.size();
}

std::string json_str(const std::string &s) {
  std::ostringstream o;
  o << '"';
  for (char c : s) {
    if (c == '"' || c == '\\')
      o << '\\' << c;
    else if (c == '\n')
      o << "\\n";
    else if ((unsigned char)c < 0x20)
      o << "\\u" << std::hex << std::setw(4) << std::setfill('0') << (int)c << std::dec;
    else
      o << c;
  }
  o << '"';
  return o.str();
}

std::string params_json(const RunParams &p, const std::string &warnings) {
  std::ostringstream o;
  o << std::setprecision(17);
  o << "{";
  auto num = [&](const char *k, double v) { o << json_str(k) << ":" << v << ","; };
  auto integer = [&](const char *k, long long v) { o << json_str(k) << ":" << v << ","; };
  auto str = [&](const char *k, const std::string &v) { o << json_str(k) << ":" << json_str(v) << ","; };
  num("tend", p.tend);
  num("save_freq", p.save_freq);
  integer("multiple_outputs", p.multiple_outputs);
  str("restart_file", p.restart_file);
  str("output_filename", p.filename_out);
  str("output_path", p.output_path);
  integer("time_stepping", p.time_stepping);
  str("problem", p.problem);
  integer("seed", p.seed);
  integer("log_frequency", p.log_frequency);
  num("epsilon_reset_negative", p.epsilon_reset_negative);
  num("epsilon", p.epsilon);
  integer("Nx", p.Nx);
  integer("Ny", p.Ny);
  integer("Ng", p.Ng);
  num("xmin", p.xmin);
  num("xmax", p.xmax);
  num("ymin", p.ymin);
  num("ymax", p.ymax);
  num("dx", p.dx);
  num("dy", p.dy);
  num("CFL", p.CFL);
  integer("boundary_x", p.boundary_x);
  integer("boundary_y", p.boundary_y);
  integer("reconstruction", p.reconstruction);
  integer("riemann_solver", p.riemann_solver);
  num("gamma0", p.gamma0);
  num("m1", p.m1);
  num("theta1", p.theta1);
  integer("well_balanced_flux_at_y_bc", p.well_balanced_flux_at_y_bc);
  num("fslp_K", p.fslp_K);
  integer("gravity_mode", p.gravity_mode);
  num("gx", p.gx);
  num("gy", p.gy);
  integer("tc_active", p.thermal_conductivity_active);
  integer("tc_mode", p.thermal_conductivity_mode);
  num("kappa", p.kappa);
  integer("bctc_ymin", p.bctc_ymin);
  integer("bctc_ymax", p.bctc_ymax);
  num("bctc_ymin_value", p.bctc_ymin_value);
  num("bctc_ymax_value", p.bctc_ymax_value);
  integer("visc_active", p.viscosity_active);
  num("mu", p.mu);
  num("hot_bubble_g0", p.hot_bubble_g0);
  num("kh_a", p.kh_a);
  num("kh_amp", p.kh_amp);
  num("kh_P0", p.kh_P0);
  num("kh_rho_fac", p.kh_rho_fac);
  num("kh_sigma", p.kh_sigma);
  num("kh_uflow", p.kh_uflow);
  num("kh_y1", p.kh_y1);
  num("kh_y2", p.kh_y2);
  num("gresho_density", p.gresho_density);
  num("gresho_Mach", p.gresho_Mach);
  num("h84_pert", p.h84_pert);
  num("c91_pert", p.c91_pert);
  integer("parse_error", p.reader.ini.parse_error());
  std::ostringstream dump;
  p.reader.output_values(dump);
  str("last_ini", dump.str());
  o << json_str("warnings") << ":" << json_str(warnings) << "}";
  return o.str();
}

RunParams load(const char *path, int is_text, std::ostream &warn) {
  return is_text ? read_ini_text(path, warn) : read_ini(path, warn);
}

}  // namespace

extern "C" {

const char *fv2dh_last_error(void) { return g_err.c_str(); }

// JSON of the parameters read from a .ini file (is_text: `src` is the file's text),
// including "warnings" and "last_ini". Returns the length (copy truncated to cap), or -1.
int fv2dh_read_ini(const char *src, int is_text, char *out, int cap) {
  try {
    std::ostringstream warn;
    const RunParams p = load(src, is_text, warn);
    return copy_out(params_json(p, warn.str()), out, cap);
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// fv2d_params of the run (precision as given). 0 or -1.
int fv2dh_fv2d_params(const char *src, int is_text, int precision, fv2d_params *out) {
  try {
    std::ostringstream warn;
    *out = to_fv2d_params(load(src, is_text, warn), precision);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// Initial primitives of the run's problem into q (4*Nx*Ny). 0 or -1.
int fv2dh_init_problem(const char *src, int is_text, double *q) {
  try {
    std::ostringstream warn;
    init_problem(load(src, is_text, warn), q);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// IOManager of the run (creates the output directory, writes last.ini in the current
// directory and the .ini copy), then one snapshot of q. 0 or -1.
int fv2dh_save(const char *src, int is_text, const double *q, int iteration, double t, int force_truncate) {
  try {
    std::ostringstream warn, out;
    const RunParams p = load(src, is_text, warn);
    IOManager io(p, out);
    io.force_file_truncation = force_truncate != 0;
    io.save_solution(q, iteration, t);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// Restart state of the run into q (4*Nx*Ny); time and iteration. 0 or -1 (message in
// fv2dh_last_error).
int fv2dh_load(const char *src, int is_text, double *q, double *time, int *iteration) {
  try {
    std::ostringstream warn, out;
    const RunParams p = load(src, is_text, warn);
    IOManager io(p, out);
    std::vector<double> v;
    const RestartInfo r = io.load_snapshot(v);
    std::memcpy(q, v.data(), v.size() * sizeof(double));
    *time = r.time;
    *iteration = r.iteration;
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// Dataset `name` of an HDF5 file: number of values (copied up to cap), or -1.
long fv2dh_h5_read(const char *file, const char *name, double *out, long cap) {
  try {
    hid_t f = h5_open(file, false);
    std::vector<double> v;
    try {
      v = h5_read_vector(f, name);
    } catch (...) {
      h5_close(f);
      throw;
    }
    h5_close(f);
    if (out) std::memcpy(out, v.data(), sizeof(double) * std::min((long)v.size(), cap));
    return (long)v.size();
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// Attribute `name` of object `obj` ("/" or a group) as text: kind 0 int, 1 double (17
// digits), 2 string. Length or -1.
int fv2dh_h5_attr(const char *file, const char *obj, const char *name, int kind, char *out, int cap) {
  try {
    hid_t f = h5_open(file, false);
    std::string s;
    try {
      hid_t g = (std::strcmp(obj, "/") == 0) ? f : h5_group_open(f, obj);
      std::ostringstream o;
      o << std::setprecision(17);
      if (kind == 0)
        o << h5_read_attr_int(g, name);
      else if (kind == 1)
        o << h5_read_attr_double(g, name);
      else
        o << h5_read_attr_string(g, name);
      s = o.str();
      if (g != f) h5_group_close(g);
    } catch (...) {
      h5_close(f);
      throw;
    }
    h5_close(f);
    return copy_out(s, out, cap);
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// Names of the objects of a group, one per line, in name order. Length or -1.
int fv2dh_h5_list(const char *file, const char *group, char *out, int cap) {
  try {
    hid_t f = h5_open(file, false);
    std::string s;
    try {
      hid_t g = (std::strcmp(group, "/") == 0) ? f : h5_group_open(f, group);
      for (const std::string &n : h5_list(g)) s += n + "\n";
      if (g != f) h5_group_close(g);
    } catch (...) {
      h5_close(f);
      throw;
    }
    h5_close(f);
    return copy_out(s, out, cap);
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

}  // extern "C"
