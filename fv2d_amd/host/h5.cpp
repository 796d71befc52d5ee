// h5.cpp -- run-time loaded HDF5 C API and small helpers (see h5.h).
#include "h5.h"

#include <dlfcn.h>

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>

namespace fv2dh {

namespace {

template <typename F>
void bind(void *lib, F &fn, const char *name) {
  void *s = dlsym(lib, name);
  if (!s) throw std::runtime_error(std::string("libhdf5: missing symbol ") + name);
  fn = reinterpret_cast<F>(s);
}

hid_t global_id(void *lib, const char *name) {
  void *s = dlsym(lib, name);
  if (!s) throw std::runtime_error(std::string("libhdf5: missing symbol ") + name);
  return *reinterpret_cast<hid_t *>(s);
}

H5 load() {
  const char *env = std::getenv("FV2D_HDF5_LIB");
  const char *cands[] = {env, "/opt/conda/lib/libhdf5.so.103", "libhdf5.so.103", "libhdf5.so"};
  void *lib = nullptr;
  std::string tried;
  for (const char *c : cands) {
    if (!c || !*c) continue;
    lib = dlopen(c, RTLD_NOW | RTLD_LOCAL);
    if (lib) break;
    tried += std::string(" ") + c;
  }
  if (!lib) throw std::runtime_error("cannot load libhdf5 (tried" + tried + "); set FV2D_HDF5_LIB");
  H5 a;
  bind(lib, a.open, "H5open");
  bind(lib, a.eset_auto, "H5Eset_auto2");
  bind(lib, a.fcreate, "H5Fcreate");
  bind(lib, a.fopen, "H5Fopen");
  bind(lib, a.fclose, "H5Fclose");
  bind(lib, a.gcreate, "H5Gcreate2");
  bind(lib, a.gopen, "H5Gopen2");
  bind(lib, a.gclose, "H5Gclose");
  bind(lib, a.gget_info, "H5Gget_info");
  bind(lib, a.lget_name_by_idx, "H5Lget_name_by_idx");
  bind(lib, a.lexists, "H5Lexists");
  bind(lib, a.screate, "H5Screate");
  bind(lib, a.screate_simple, "H5Screate_simple");
  bind(lib, a.sclose, "H5Sclose");
  bind(lib, a.sget_ndims, "H5Sget_simple_extent_ndims");
  bind(lib, a.sget_dims, "H5Sget_simple_extent_dims");
  bind(lib, a.dcreate, "H5Dcreate2");
  bind(lib, a.dopen, "H5Dopen2");
  bind(lib, a.dwrite, "H5Dwrite");
  bind(lib, a.dread, "H5Dread");
  bind(lib, a.dclose, "H5Dclose");
  bind(lib, a.dget_space, "H5Dget_space");
  bind(lib, a.acreate, "H5Acreate2");
  bind(lib, a.aopen, "H5Aopen");
  bind(lib, a.aexists, "H5Aexists");
  bind(lib, a.awrite, "H5Awrite");
  bind(lib, a.aread, "H5Aread");
  bind(lib, a.aclose, "H5Aclose");
  bind(lib, a.aget_type, "H5Aget_type");
  bind(lib, a.tcopy, "H5Tcopy");
  bind(lib, a.tset_size, "H5Tset_size");
  bind(lib, a.tset_cset, "H5Tset_cset");
  bind(lib, a.tclose, "H5Tclose");
  bind(lib, a.tis_variable_str, "H5Tis_variable_str");
  bind(lib, a.tget_size, "H5Tget_size");
  bind(lib, a.free_memory, "H5free_memory");
  if (a.open() < 0) throw std::runtime_error("H5open failed");
  a.T_NATIVE_INT = global_id(lib, "H5T_NATIVE_INT_g");
  a.T_NATIVE_DOUBLE = global_id(lib, "H5T_NATIVE_DOUBLE_g");
  a.T_C_S1 = global_id(lib, "H5T_C_S1_g");
  a.eset_auto(H5E_DEFAULT, nullptr, nullptr);  // errors are reported by the callers
  return a;
}

// File access flags (H5Fpublic.h); hdf5.h's own macros call H5open() and H5check_version(),
// which this library resolves at run time instead of linking.
constexpr unsigned kAccRdonly = 0x0000u, kAccRdwr = 0x0001u, kAccTrunc = 0x0002u;

void check(bool ok, const std::string &what) {
  if (!ok) throw std::runtime_error("HDF5: " + what);
}

}  // namespace

const H5 &h5() {
  static std::once_flag once;
  static H5 api;
  static std::string err;
  std::call_once(once, [] {
    try {
      api = load();
    } catch (const std::exception &e) {
      err = e.what();
    }
  });
  if (!err.empty()) throw std::runtime_error(err);
  return api;
}

hid_t h5_create(const std::string &path) {
  hid_t f = h5().fcreate(path.c_str(), kAccTrunc, H5P_DEFAULT, H5P_DEFAULT);
  check(f >= 0, "cannot create " + path);
  return f;
}

hid_t h5_open(const std::string &path, bool write) {
  hid_t f = h5().fopen(path.c_str(), write ? kAccRdwr : kAccRdonly, H5P_DEFAULT);
  check(f >= 0, "cannot open " + path);
  return f;
}

void h5_close(hid_t f) { h5().fclose(f); }

static void write_attr(hid_t obj, const char *name, hid_t type, const void *v) {
  const H5 &a = h5();
  hid_t sp = a.screate(H5S_SCALAR);
  hid_t at = a.acreate(obj, name, type, sp, H5P_DEFAULT, H5P_DEFAULT);
  check(at >= 0, std::string("cannot create attribute ") + name);
  herr_t e = a.awrite(at, type, v);
  a.aclose(at);
  a.sclose(sp);
  check(e >= 0, std::string("cannot write attribute ") + name);
}

void h5_attr_int(hid_t obj, const char *name, int v) { write_attr(obj, name, h5().T_NATIVE_INT, &v); }

void h5_attr_double(hid_t obj, const char *name, double v) { write_attr(obj, name, h5().T_NATIVE_DOUBLE, &v); }

void h5_attr_string(hid_t obj, const char *name, const std::string &v) {
  // variable-length string, ASCII, scalar
  const H5 &a = h5();
  hid_t t = a.tcopy(a.T_C_S1);
  a.tset_size(t, H5T_VARIABLE);
  a.tset_cset(t, H5T_CSET_ASCII);
  const char *p = v.c_str();
  write_attr(obj, name, t, &p);
  a.tclose(t);
}

bool h5_has_attr(hid_t obj, const char *name) { return h5().aexists(obj, name) > 0; }

static void read_attr(hid_t obj, const char *name, hid_t type, void *v) {
  const H5 &a = h5();
  hid_t at = a.aopen(obj, name, H5P_DEFAULT);
  check(at >= 0, std::string("no attribute ") + name);
  herr_t e = a.aread(at, type, v);
  a.aclose(at);
  check(e >= 0, std::string("cannot read attribute ") + name);
}

int h5_read_attr_int(hid_t obj, const char *name) {
  int v = 0;
  read_attr(obj, name, h5().T_NATIVE_INT, &v);
  return v;
}

double h5_read_attr_double(hid_t obj, const char *name) {
  double v = 0;
  read_attr(obj, name, h5().T_NATIVE_DOUBLE, &v);
  return v;
}

std::string h5_read_attr_string(hid_t obj, const char *name) {
  const H5 &a = h5();
  hid_t at = a.aopen(obj, name, H5P_DEFAULT);
  check(at >= 0, std::string("no attribute ") + name);
  hid_t ft = a.aget_type(at);
  std::string out;
  if (a.tis_variable_str(ft) > 0) {
    hid_t t = a.tcopy(a.T_C_S1);
    a.tset_size(t, H5T_VARIABLE);
    char *p = nullptr;
    herr_t e = a.aread(at, t, &p);
    if (e >= 0 && p) {
      out = p;
      a.free_memory(p);
    }
    a.tclose(t);
  } else {
    const size_t n = a.tget_size(ft);
    std::vector<char> buf(n + 1, 0);
    hid_t t = a.tcopy(a.T_C_S1);
    a.tset_size(t, n);
    a.aread(at, t, buf.data());
    a.tclose(t);
    out = buf.data();
  }
  a.tclose(ft);
  a.aclose(at);
  return out;
}

void h5_write_vector(hid_t loc, const char *name, const double *v, size_t n) {
  const H5 &a = h5();
  hsize_t dims[1] = {(hsize_t)n};
  hid_t sp = a.screate_simple(1, dims, nullptr);
  hid_t ds = a.dcreate(loc, name, a.T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  check(ds >= 0, std::string("cannot create dataset ") + name);
  herr_t e = a.dwrite(ds, a.T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v);
  a.dclose(ds);
  a.sclose(sp);
  check(e >= 0, std::string("cannot write dataset ") + name);
}

size_t h5_vector_size(hid_t loc, const std::string &name) {
  const H5 &a = h5();
  hid_t ds = a.dopen(loc, name.c_str(), H5P_DEFAULT);
  check(ds >= 0, "no dataset " + name);
  hid_t sp = a.dget_space(ds);
  hsize_t dims[8] = {0};
  const int nd = a.sget_ndims(sp);
  check(nd >= 1 && nd <= 8, "bad rank of " + name);
  a.sget_dims(sp, dims, nullptr);
  a.sclose(sp);
  a.dclose(ds);
  return (size_t)dims[0];
}

std::vector<double> h5_read_vector(hid_t loc, const std::string &name) {
  const H5 &a = h5();
  hid_t ds = a.dopen(loc, name.c_str(), H5P_DEFAULT);
  check(ds >= 0, "no dataset " + name);
  hid_t sp = a.dget_space(ds);
  hsize_t dims[8] = {0};
  const int nd = a.sget_ndims(sp);
  a.sget_dims(sp, dims, nullptr);
  a.sclose(sp);
  size_t n = 1;
  for (int k = 0; k < nd; ++k) n *= (size_t)dims[k];
  std::vector<double> v(n);
  herr_t e = a.dread(ds, a.T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
  a.dclose(ds);
  check(e >= 0, "cannot read dataset " + name);
  return v;
}

hid_t h5_group_create(hid_t loc, const std::string &name) {
  hid_t g = h5().gcreate(loc, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  check(g >= 0, "cannot create group " + name);
  return g;
}

hid_t h5_group_open(hid_t loc, const std::string &name) {
  hid_t g = h5().gopen(loc, name.c_str(), H5P_DEFAULT);
  check(g >= 0, "no group " + name);
  return g;
}

void h5_group_close(hid_t g) { h5().gclose(g); }

std::vector<std::string> h5_list(hid_t loc) {
  const H5 &a = h5();
  H5G_info_t info;
  check(a.gget_info(loc, &info) >= 0, "cannot list group");
  std::vector<std::string> names;
  for (hsize_t k = 0; k < info.nlinks; ++k) {
    char buf[512];
    ssize_t n = a.lget_name_by_idx(loc, ".", H5_INDEX_NAME, H5_ITER_INC, k, buf, sizeof buf, H5P_DEFAULT);
    if (n >= 0) names.emplace_back(buf, (size_t)std::min<ssize_t>(n, (ssize_t)sizeof buf - 1));
  }
  return names;
}

}  // namespace fv2dh
