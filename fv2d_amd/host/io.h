// io.h -- snapshots and restart with the reference's HDF5 layout (IOManager.h:76-398).
//
// Layout (written with the HDF5 C API):
//  - unique mode (one file <output_path>/<output_filename>.h5 for the run): root attributes
//    Ntx, Nty, Nx, Ny, ibeg, iend, jbeg, jend (int) and problem (string), root datasets x
//    and y (vertex coordinates, (Nx+1)*(Ny+1), j-outer), and per snapshot a group
//    "ite_%04d" holding rho, u, v, prs (Nx*Ny doubles, j-outer) with attributes time
//    (double) and iteration (int). An XDMF sidecar <output_filename>.xmf lists every
//    snapshot as a grid of one temporal collection.
//  - multiple mode (one file <output_filename>_%04d.h5 per snapshot): the same root
//    attributes and x, y, then rho, u, v, prs and attributes time, iteration at the root,
//    with an XDMF file per snapshot.
// Restart (loadSnapshot, IOManager.h:284-398): "file.h5" (last snapshot, or the root of a
// multiple-mode file) or "file.h5:/ite_0005"; same resolution required; restart time must
// be before tend; restarting from the run's own unique file continues it, anything else
// truncates the output and writes the loaded state as the first snapshot.
#pragma once

#include <string>
#include <vector>

#include "params.h"

namespace fv2dh {

struct RestartInfo {
  double time;
  int iteration;
};

class IOManager {
 public:
  // Creates output_path if needed (with a message on out), writes last.ini and
  // <output_path>/<output_filename>.ini.
  IOManager(const RunParams &p, std::ostream &out);

  // q: domain primitives, 4*Nx*Ny, snapshot order.
  void save_solution(const double *q, int iteration, double t);
  // Loads the restart state into q (4*Nx*Ny); throws std::runtime_error on the reference's
  // error cases.
  RestartInfo load_snapshot(std::vector<double> &q);

  bool force_file_truncation = false;

 private:
  void save_multiple(const double *q, int iteration, double t);
  void save_unique(const double *q, int iteration, double t);
  void write_mesh(long long file) const;
  const RunParams &p_;
  std::ostream &out_;
};

// XDMF text of the sidecar files (exposed for tests).
std::string xdmf_header(const RunParams &p, const std::string &h5name);
std::string xdmf_grid(const std::string &name, double time, const std::string &group);
std::string xdmf_footer();

}  // namespace fv2dh
