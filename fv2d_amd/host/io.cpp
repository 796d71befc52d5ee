// io.cpp -- HDF5 snapshots, XDMF sidecars and restart (see io.h).
#include "io.h"

#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "h5.h"

namespace fv2dh {

namespace fs = std::filesystem;

namespace {

std::string ite_name(int iteration) {
  std::ostringstream o;
  o << "ite_" << std::setw(4) << std::setfill('0') << iteration;
  return o.str();
}

void split_fields(const RunParams &p, const double *q, std::vector<double> (&f)[4]) {
  const size_t n = (size_t)p.Nx * p.Ny;
  for (int k = 0; k < 4; ++k) f[k].assign(q + k * n, q + (k + 1) * n);
}

const char *kFields[4] = {"rho", "u", "v", "prs"};

}  // namespace

// ------------------------------------------------------------------ XDMF

std::string xdmf_header(const RunParams &p, const std::string &h5name) {
  std::ostringstream o;
  o << "<?xml version=\"1.0\" ?>\n"
    << "<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" [\n"
    << "<!ENTITY file \"" << h5name << ":\">\n"
    << "<!ENTITY fdim \"" << p.Ny << " " << p.Nx << "\">\n"
    << "<!ENTITY gdim \"" << p.Ny + 1 << " " << p.Nx + 1 << "\">\n"
    << "<!ENTITY GridEntity '\n"
    << "<Topology TopologyType=\"2DSMesh\" Dimensions=\"&gdim;\"/>\n"
    << "<Geometry GeometryType=\"X_Y\">\n"
    << "  <DataItem Dimensions=\"&gdim;\" NumberType=\"Float\" Precision=\"8\" Format=\"HDF\">&file;/x</DataItem>\n"
    << "  <DataItem Dimensions=\"&gdim;\" NumberType=\"Float\" Precision=\"8\" Format=\"HDF\">&file;/y</DataItem>\n"
    << "</Geometry>'>\n"
    << "]>\n"
    << "<Xdmf Version=\"3.0\">\n"
    << "<Domain>\n"
    << "  <Grid Name=\"TimeSeries\" GridType=\"Collection\" CollectionType=\"Temporal\">";
  return o.str();
}

std::string xdmf_grid(const std::string &name, double time, const std::string &group) {
  char tbuf[64];
  snprintf(tbuf, sizeof tbuf, "%lf", time);
  auto item = [&](const char *field) {
    return std::string("<DataItem Dimensions=\"&fdim;\" NumberType=\"Float\" Precision=\"8\" Format=\"HDF\">&file;/") +
           group + field + "</DataItem>";
  };
  std::ostringstream o;
  o << "\n    <Grid Name=\"" << name << "\" GridType=\"Uniform\">\n"
    << "      <Time Value=\"" << tbuf << "\" />\n"
    << "      &GridEntity;\n"
    << "      <Attribute Name=\"rho\" AttributeType=\"Scalar\" Center=\"Cell\">\n"
    << "        " << item("rho") << "\n"
    << "      </Attribute>\n"
    << "      <Attribute Name=\"velocity\" AttributeType=\"Vector\" Center=\"Cell\">\n"
    << "        <DataItem Dimensions=\"&fdim; 2\" ItemType=\"Function\" Function=\"JOIN($0, $1)\">\n"
    << "          " << item("u") << "\n"
    << "          " << item("v") << "\n"
    << "        </DataItem>\n"
    << "      </Attribute>\n"
    << "      <Attribute Name=\"prs\" AttributeType=\"Scalar\" Center=\"Cell\">\n"
    << "        " << item("prs") << "\n"
    << "      </Attribute>\n"
    << "    </Grid>";
  return o.str();
}

std::string xdmf_footer() { return "\n  </Grid>\n</Domain>\n</Xdmf>\n"; }

// ------------------------------------------------------------------ IOManager

IOManager::IOManager(const RunParams &p, std::ostream &out) : p_(p), out_(out) {
  if (!fs::exists(p.output_path)) {
    out_ << "Output path does not exist, creating directory `" << p.output_path << "`." << std::endl;
    fs::create_directory(p.output_path);
  }
  std::ofstream local("last.ini");
  p.reader.output_values(local);
  std::ofstream copy(p.output_path + "/" + p.filename_out + ".ini");
  p.reader.output_values(copy);
}

void IOManager::save_solution(const double *q, int iteration, double t) {
  if (p_.multiple_outputs)
    save_multiple(q, iteration, t);
  else
    save_unique(q, iteration, t);
}

void IOManager::write_mesh(long long file) const {
  const hid_t f = (hid_t)file;
  h5_attr_int(f, "Ntx", p_.Ntx);
  h5_attr_int(f, "Nty", p_.Nty);
  h5_attr_int(f, "Nx", p_.Nx);
  h5_attr_int(f, "Ny", p_.Ny);
  h5_attr_int(f, "ibeg", p_.ibeg);
  h5_attr_int(f, "iend", p_.iend);
  h5_attr_int(f, "jbeg", p_.jbeg);
  h5_attr_int(f, "jend", p_.jend);
  h5_attr_string(f, "problem", p_.problem);
  std::vector<double> x, y;
  x.reserve((size_t)(p_.Nx + 1) * (p_.Ny + 1));
  y.reserve(x.capacity());
  for (int j = p_.jbeg; j <= p_.jend; ++j)
    for (int i = p_.ibeg; i <= p_.iend; ++i) {
      x.push_back((i - p_.ibeg) * p_.dx + p_.xmin);
      y.push_back((j - p_.jbeg) * p_.dy + p_.ymin);
    }
  h5_write_vector(f, "x", x.data(), x.size());
  h5_write_vector(f, "y", y.data(), y.size());
}

void IOManager::save_multiple(const double *q, int iteration, double t) {
  std::ostringstream o;
  o << p_.filename_out << "_" << std::setw(4) << std::setfill('0') << iteration;
  const std::string stem = o.str(), h5name = stem + ".h5", xmfname = stem + ".xmf";
  const std::string dir = p_.output_path + "/";
  hid_t f = h5_create(dir + h5name);
  FILE *x = fopen((dir + xmfname).c_str(), "w+");
  if (!x) {
    h5_close(f);
    throw std::runtime_error("Failed to open XDMF file for writing: " + dir + xmfname);
  }
  write_mesh(f);
  std::vector<double> fld[4];
  split_fields(p_, q, fld);
  for (int k = 0; k < 4; ++k) h5_write_vector(f, kFields[k], fld[k].data(), fld[k].size());
  h5_attr_double(f, "time", t);
  h5_attr_int(f, "iteration", iteration);
  h5_close(f);
  const std::string doc = xdmf_header(p_, h5name) + xdmf_grid(stem, t, "") + xdmf_footer();
  fwrite(doc.data(), 1, doc.size(), x);
  fclose(x);
}

void IOManager::save_unique(const double *q, int iteration, double t) {
  const std::string name = ite_name(iteration);
  const std::string h5name = p_.filename_out + ".h5", xmfname = p_.filename_out + ".xmf";
  const std::string dir = p_.output_path + "/";
  force_file_truncation = force_file_truncation || iteration == 0;
  const bool trunc = force_file_truncation;
  hid_t f = trunc ? h5_create(dir + h5name) : h5_open(dir + h5name, true);
  FILE *x = fopen((dir + xmfname).c_str(), trunc ? "w+" : "r+");
  if (!x) {
    h5_close(f);
    throw std::runtime_error("Failed to open XDMF file '" + dir + xmfname + "' with mode '" + (trunc ? "w+" : "r+") +
                             "'.");
  }
  if (trunc) {
    force_file_truncation = false;
    write_mesh(f);
    const std::string head = xdmf_header(p_, h5name) + xdmf_footer();
    fwrite(head.data(), 1, head.size(), x);
  }
  std::vector<double> fld[4];
  split_fields(p_, q, fld);
  hid_t g = h5_group_create(f, name);
  for (int k = 0; k < 4; ++k) h5_write_vector(g, kFields[k], fld[k].data(), fld[k].size());
  h5_attr_double(g, "time", t);
  h5_attr_int(g, "iteration", iteration);
  h5_group_close(g);
  h5_close(f);
  // replace the footer by this snapshot's grid and the footer
  const std::string foot = xdmf_footer();
  fseek(x, -(long)foot.size(), SEEK_END);
  const std::string tail = xdmf_grid(name, t, name + "/") + foot;
  fwrite(tail.data(), 1, tail.size(), x);
  fclose(x);
}

RestartInfo IOManager::load_snapshot(std::vector<double> &q) {
  std::string file = p_.restart_file, group;
  const size_t d = file.find(".h5:/");
  if (d != std::string::npos) {
    group = file.substr(d + 5);
    file.resize(d + 3);
  }
  const std::string own = p_.output_path + "/" + p_.filename_out + ".h5";
  if (!p_.multiple_outputs && fs::equivalent(file, own)) {
    if (d != std::string::npos) {
      std::cerr << "Invalid restart file : if your restart file and output file are "
                   "the same, you can only start from the last iteration."
                << std::endl
                << std::endl;
      throw std::runtime_error("ERROR : Invalid restart_file.");
    }
  } else {
    force_file_truncation = true;
  }
  hid_t f = h5_open(file, false);
  double time = 0.0;
  int iteration = 0;
  try {
    if (h5_has_attr(f, "time")) {
      time = h5_read_attr_double(f, "time");
      iteration = h5_read_attr_int(f, "iteration");
    } else {
      if (group.empty()) {
        // objects in name order: ite_0000 .. ite_NNNN, x, y -> the last snapshot group
        const std::vector<std::string> names = h5_list(f);
        if (names.size() < 3) throw std::runtime_error("ERROR : no snapshot in " + file);
        group = names[names.size() - 3];
      }
      hid_t g = h5_group_open(f, group);
      time = h5_read_attr_double(g, "time");
      iteration = h5_read_attr_int(g, "iteration");
      h5_group_close(g);
      group += "/";
    }
    const size_t nt = h5_vector_size(f, group + "rho");
    if (nt != (size_t)p_.Nx * p_.Ny) {
      std::cerr << "Attempting to restart with a different resolution ! Ncells (restart) = " << nt
                << "; Run resolution = " << p_.Nx << "x" << p_.Ny << "=" << p_.Nx * p_.Ny << std::endl;
      throw std::runtime_error("ERROR : Trying to restart from a file with a different resolution !");
    }
    out_ << "Loading restart data from hdf5" << std::endl;
    const size_t n = (size_t)p_.Nx * p_.Ny;
    q.assign(4 * n, 0.0);
    for (int k = 0; k < 4; ++k) {
      const std::vector<double> v = h5_read_vector(f, group + kFields[k]);
      std::copy(v.begin(), v.begin() + n, q.begin() + k * n);
    }
  } catch (...) {
    h5_close(f);
    throw;
  }
  h5_close(f);
  if (time + p_.epsilon > p_.tend) {
    std::cerr << "Restart time is greater than end time : " << std::endl
              << "  time: " << time << "\ttend: " << p_.tend << std::endl
              << std::endl;
    throw std::runtime_error("ERROR : restart time is greater than the end time.");
  }
  out_ << "Restart finished !" << std::endl;
  if (force_file_truncation) save_solution(q.data(), iteration, time);
  return {time, iteration};
}

}  // namespace fv2dh
