// h5.h -- the few HDF5 C API calls the snapshot writer/reader needs, resolved at run time
// from libhdf5 (dlopen), so that neither the solver executable nor a Python process that
// loads the host library links against the HDF5 installation's other libraries.
//
// Library search: $FV2D_HDF5_LIB, then /opt/conda/lib/libhdf5.so.103, then libhdf5.so on
// the loader's path. Types and constants come from hdf5.h (declarations only).
#pragma once

#include <hdf5.h>

#include <string>
#include <vector>

namespace fv2dh {

struct H5 {
  decltype(&H5open) open;
  decltype(&H5Eset_auto2) eset_auto;
  decltype(&H5Fcreate) fcreate;
  decltype(&H5Fopen) fopen;
  decltype(&H5Fclose) fclose;
  decltype(&H5Gcreate2) gcreate;
  decltype(&H5Gopen2) gopen;
  decltype(&H5Gclose) gclose;
  decltype(&H5Gget_info) gget_info;
  decltype(&H5Lget_name_by_idx) lget_name_by_idx;
  decltype(&H5Lexists) lexists;
  decltype(&H5Screate) screate;
  decltype(&H5Screate_simple) screate_simple;
  decltype(&H5Sclose) sclose;
  decltype(&H5Sget_simple_extent_ndims) sget_ndims;
  decltype(&H5Sget_simple_extent_dims) sget_dims;
  decltype(&H5Dcreate2) dcreate;
  decltype(&H5Dopen2) dopen;
  decltype(&H5Dwrite) dwrite;
  decltype(&H5Dread) dread;
  decltype(&H5Dclose) dclose;
  decltype(&H5Dget_space) dget_space;
  decltype(&H5Acreate2) acreate;
  decltype(&H5Aopen) aopen;
  decltype(&H5Aexists) aexists;
  decltype(&H5Awrite) awrite;
  decltype(&H5Aread) aread;
  decltype(&H5Aclose) aclose;
  decltype(&H5Aget_type) aget_type;
  decltype(&H5Tcopy) tcopy;
  decltype(&H5Tset_size) tset_size;
  decltype(&H5Tset_cset) tset_cset;
  decltype(&H5Tclose) tclose;
  decltype(&H5Tis_variable_str) tis_variable_str;
  decltype(&H5Tget_size) tget_size;
  decltype(&H5free_memory) free_memory;
  hid_t T_NATIVE_INT, T_NATIVE_DOUBLE, T_C_S1;
};

// The loaded API (throws std::runtime_error when libhdf5 cannot be loaded).
const H5 &h5();

// Thin RAII-free helpers over H5 (each throws std::runtime_error with the object name on
// failure). Files are identified by hid_t; paths are absolute within the file.
hid_t h5_create(const std::string &path);  // truncate
hid_t h5_open(const std::string &path, bool write);
void h5_close(hid_t f);
void h5_attr_int(hid_t obj, const char *name, int v);
void h5_attr_double(hid_t obj, const char *name, double v);
void h5_attr_string(hid_t obj, const char *name, const std::string &v);
bool h5_has_attr(hid_t obj, const char *name);
int h5_read_attr_int(hid_t obj, const char *name);
double h5_read_attr_double(hid_t obj, const char *name);
std::string h5_read_attr_string(hid_t obj, const char *name);
void h5_write_vector(hid_t loc, const char *name, const double *v, size_t n);
std::vector<double> h5_read_vector(hid_t loc, const std::string &name);
size_t h5_vector_size(hid_t loc, const std::string &name);
hid_t h5_group_create(hid_t loc, const std::string &name);
hid_t h5_group_open(hid_t loc, const std::string &name);
void h5_group_close(hid_t g);
// Names of the objects of a group in name order (as HDF5's default link index).
std::vector<std::string> h5_list(hid_t loc);

}  // namespace fv2dh
