// driver.cpp -- main.cpp's time loop over the C ABI (see driver.h).
#include "driver.h"

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fv2d_amd.h"
#include "init.h"
#include "io.h"

namespace fv2dh {

namespace {

struct Handle {
  fv2d_handle *h = nullptr;
  ~Handle() { fv2d_destroy(h); }
  void ok(int rc, const char *what) const {
    if (rc) throw std::runtime_error(std::string(what) + ": " + fv2d_last_error(h));
  }
};

// checkNegatives' report (SimInfo.h:637-645)
void report(const fv2d_step_stats &s, std::ostream &out) {
  if (s.negative_density) out << "--> negative density: " << s.negative_density << std::endl;
  if (s.negative_pressure) out << "--> negative pressure: " << s.negative_pressure << std::endl;
  if (s.nan_count) out << "--> NaN detected." << std::endl;
}

// ComputeDtFunctor's diagnostic line (ComputeDt.h:55-62)
void diag_line(const RunParams &p, double t, const double inv[3], std::ostream &out) {
  out << "Computing dts at (t=" << t << ") : dt_hyp=" << 1.0 / inv[0];
  if (p.thermal_conductivity_active) out << "; dt_TC=" << 1.0 / inv[1];
  if (p.viscosity_active) out << "; dt_visc=" << 1.0 / inv[2];
  out << std::endl;
}

}  // namespace

int run_simulation(const RunParams &p, const DriverOptions &opt, std::ostream &out) {
  const fv2d_params fp = to_fv2d_params(p, opt.precision);
  Handle H;
  if (int rc = fv2d_create(&fp, &H.h)) throw std::runtime_error(std::string("fv2d_create: ") + fv2d_last_error(nullptr) +
                                                                 " (code " + std::to_string(rc) + ")");
  double t = 0.0;
  int ite = 0;
  double next_save = 0.0;
  IOManager io(p, out);
  std::vector<double> q;
  if (!p.restart_file.empty()) {
    const RestartInfo r = io.load_snapshot(q);
    t = r.time;
    ite = r.iteration;
    out << "Restart at iteration " << ite << " and time " << t << std::endl;
    next_save = t + p.save_freq;
    ite++;
  } else {
    q = init_problem(p);
  }
  H.ok(fv2d_set_prim(H.h, q.data()), "fv2d_set_prim");

  const double eps = p.epsilon;
  int next_log = 0;
  while (t + eps < p.tend) {
    const bool save_needed = t + eps > next_save;
    const bool diag = next_log == 0;
    if (opt.per_step || diag || save_needed) {
      double dt = 0.0, inv[3];
      H.ok(fv2d_compute_dt(H.h, &dt, inv), "fv2d_compute_dt");
      if (diag) diag_line(p, t, inv, out);
      next_log = diag ? p.log_frequency : next_log - 1;
      if (save_needed) {
        out << " - Saving at time " << t << std::endl;
        H.ok(fv2d_get_prim(H.h, q.data()), "fv2d_get_prim");
        io.save_solution(q.data(), ite++, t);
        next_save += p.save_freq;
      }
      fv2d_step_stats s;
      H.ok(fv2d_step(H.h, dt, &s), "fv2d_step");
      report(s, out);
      t += dt;
      continue;
    }
    // next_log > 0 plain iterations follow before the next log line; they run on the
    // device until then, or until t + epsilon reaches the next snapshot time or tend
    const double stop = std::min(p.tend, next_save);
    double t2 = t;
    int32_t n = 0;
    fv2d_step_stats s;
    H.ok(fv2d_advance(H.h, &t2, stop, next_log, nullptr, &n, &s), "fv2d_advance");
    report(s, out);
    if (n == 0) {
      // t + epsilon == next_save exactly: the loop takes one more step before saving
      double dt = 0.0, inv[3];
      H.ok(fv2d_compute_dt(H.h, &dt, inv), "fv2d_compute_dt");
      next_log -= 1;
      fv2d_step_stats s1;
      H.ok(fv2d_step(H.h, dt, &s1), "fv2d_step");
      report(s1, out);
      t += dt;
      continue;
    }
    t = t2;
    next_log -= n;
  }
  out << "Time at end is " << t << std::endl;
  H.ok(fv2d_get_prim(H.h, q.data()), "fv2d_get_prim");
  io.save_solution(q.data(), ite++, t);
  return 0;
}

}  // namespace fv2dh
