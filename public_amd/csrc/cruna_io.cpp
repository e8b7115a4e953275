// CRUNA snapshot / restart files (HDF5) for the device loop's states: the
// default store/load paths of src/io/hdf5_single.f90 and get_fname of
// src/helper/io_helper.f90 for the snapshot, subset and restart types.
#include "../../include/cruna_io.h"

#include <hdf5.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const std::string &m) {
  g_err = m;
  return -1;
}

// closes an HDF5 handle when leaving scope
struct H5Handle {
  hid_t id;
  herr_t (*close)(hid_t);
  H5Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~H5Handle() {
    if (id >= 0) close(id);
  }
  H5Handle(const H5Handle &) = delete;
  H5Handle &operator=(const H5Handle &) = delete;
};

// HDF5 prints its own error stack to stderr unless told otherwise; the
// library reports through the return codes instead
void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

// the params dataset: CRUNA_IO_NPARAMS blank-padded rows of
// CRUNA_IO_PARAM_LEN characters (h5t_fortran_s1 of that size)
int write_params(hid_t file, const char *const *params, int nparams) {
  if (nparams < 0 || nparams > CRUNA_IO_NPARAMS) return fail("too many parameter rows");
  std::vector<char> buf((size_t)CRUNA_IO_NPARAMS * CRUNA_IO_PARAM_LEN, ' ');
  for (int i = 0; i < nparams; ++i) {
    const char *s = params && params[i] ? params[i] : "";
    const size_t n = std::strlen(s);
    if (n >= CRUNA_IO_PARAM_LEN) return fail("parameter row longer than the params width");
    std::memcpy(&buf[(size_t)i * CRUNA_IO_PARAM_LEN], s, n);
  }
  H5Handle type(H5Tcopy(H5T_FORTRAN_S1), H5Tclose);
  if (type.id < 0 || H5Tset_size(type.id, CRUNA_IO_PARAM_LEN) < 0) return fail("string type");
  const hsize_t dims[1] = {CRUNA_IO_NPARAMS};
  H5Handle space(H5Screate_simple(1, dims, nullptr), H5Sclose);
  H5Handle dset(H5Dcreate2(file, "params", type.id, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                H5Dclose);
  if (space.id < 0 || dset.id < 0) return fail("cannot create dataset params");
  if (H5Dwrite(dset.id, type.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0)
    return fail("cannot write dataset params");
  return 0;
}

// the whole params dataset as rows (blank-padded)
int read_params(const char *fname, std::vector<std::string> &rows) {
  quiet();
  H5Handle file(H5Fopen(fname, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (file.id < 0) return fail(std::string("cannot open ") + fname);
  H5Handle dset(H5Dopen2(file.id, "params", H5P_DEFAULT), H5Dclose);
  if (dset.id < 0) return fail("no dataset params");
  H5Handle ftype(H5Dget_type(dset.id), H5Tclose);
  H5Handle space(H5Dget_space(dset.id), H5Sclose);
  if (H5Tget_class(ftype.id) != H5T_STRING || H5Tis_variable_str(ftype.id) > 0)
    return fail("params is not a fixed-length string dataset");
  const size_t w = H5Tget_size(ftype.id);
  hsize_t n = 0;
  if (H5Sget_simple_extent_ndims(space.id) != 1 || H5Sget_simple_extent_dims(space.id, &n, nullptr) < 0)
    return fail("params is not one-dimensional");
  H5Handle mtype(H5Tcopy(H5T_FORTRAN_S1), H5Tclose);
  H5Tset_size(mtype.id, w);
  std::vector<char> buf((size_t)n * w, ' ');
  if (H5Dread(dset.id, mtype.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0)
    return fail("cannot read params");
  rows.clear();
  for (hsize_t i = 0; i < n; ++i) {
    std::string r(&buf[(size_t)i * w], w);
    const size_t e = r.find_last_not_of(" \0", std::string::npos, 2);
    rows.push_back(e == std::string::npos ? std::string() : r.substr(0, e + 1));
  }
  return 0;
}

}  // namespace

extern "C" {

const char *cruna_io_last_error(void) { return g_err.c_str(); }

int cruna_io_fname(char *out, size_t cap, const char *type, const char *workdir, int world_block,
                   int block_image, int ns, int nt) {
  if (!out || !type) return fail("null argument");
  // io_helper.f90:11-15 formats: block I2.2, image I5.5, subset I2.2, step I7.7
  char blk[64], sub[32], step[32];
  std::snprintf(blk, sizeof blk, "_%02d_%05d", world_block, block_image);
  std::snprintf(sub, sizeof sub, "_%02d", ns);
  std::snprintf(step, sizeof step, "_%07d", nt);
  std::string t(type), name;
  if (t == "data_direct_snapshot" || t == "data_adjoint_snapshot")
    name = t + "_" + blk + sub + step;  // io_helper.f90:109-114, 155-160
  else if (t == "data_direct" || t == "data_adjoint")
    name = t + "_" + blk + sub;  // :85-89, :123-127 (the whole subset)
  else if (t == "restartfile")
    name = t;  // :337-342
  else
    return fail("file type '" + t + "' is not one of the snapshot, subset or restart types");
  const std::string full = std::string(workdir ? workdir : ".") + "/" + name + ".h5";
  if (full.size() + 1 > cap) return fail("file name buffer too small");
  std::memcpy(out, full.c_str(), full.size() + 1);
  return 0;
}

int cruna_io_store_4d(const char *fname, const double *data, const int dims[4],
                      const char *const *params, int nparams) {
  if (!fname || !data || !dims) return fail("null argument");
  for (int d = 0; d < 4; ++d)
    if (dims[d] < 1) return fail("dimensions must be positive");
  quiet();
  H5Handle file(H5Fcreate(fname, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (file.id < 0) return fail(std::string("cannot create ") + fname);
  // Fortran data(d1,d2,d3,d4) is the C array [d4][d3][d2][d1]
  const hsize_t h[4] = {(hsize_t)dims[3], (hsize_t)dims[2], (hsize_t)dims[1], (hsize_t)dims[0]};
  H5Handle space(H5Screate_simple(4, h, nullptr), H5Sclose);
  H5Handle dset(H5Dcreate2(file.id, "data", H5T_NATIVE_DOUBLE, space.id, H5P_DEFAULT, H5P_DEFAULT,
                           H5P_DEFAULT),
                H5Dclose);
  if (space.id < 0 || dset.id < 0) return fail("cannot create dataset data");
  if (H5Dwrite(dset.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    return fail("cannot write dataset data");
  return write_params(file.id, params, nparams);
}

int cruna_io_load_4d(const char *fname, double *data, const int dims[4]) {
  if (!fname || !data || !dims) return fail("null argument");
  quiet();
  H5Handle file(H5Fopen(fname, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (file.id < 0) return fail(std::string("cannot open ") + fname);
  H5Handle dset(H5Dopen2(file.id, "data", H5P_DEFAULT), H5Dclose);
  if (dset.id < 0) return fail("no dataset data");
  H5Handle space(H5Dget_space(dset.id), H5Sclose);
  hsize_t h[4] = {0, 0, 0, 0};
  if (H5Sget_simple_extent_ndims(space.id) != 4 || H5Sget_simple_extent_dims(space.id, h, nullptr) < 0)
    return fail("dataset data is not four-dimensional");
  for (int d = 0; d < 4; ++d)
    if (h[3 - d] != (hsize_t)dims[d])
      return fail("dataset data has dimensions other than the ones asked for");
  if (H5Dread(dset.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    return fail("cannot read dataset data");
  return 0;
}

int cruna_io_store_restart(const char *fname, int ns, int nt, const char *const *params,
                           int nparams) {
  if (!fname) return fail("null argument");
  quiet();
  H5Handle file(H5Fcreate(fname, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (file.id < 0) return fail(std::string("cannot create ") + fname);
  const hsize_t h[1] = {2};
  const int v[2] = {ns, nt};
  H5Handle space(H5Screate_simple(1, h, nullptr), H5Sclose);
  H5Handle dset(H5Dcreate2(file.id, "data", H5T_NATIVE_INT, space.id, H5P_DEFAULT, H5P_DEFAULT,
                           H5P_DEFAULT),
                H5Dclose);
  if (space.id < 0 || dset.id < 0) return fail("cannot create dataset data");
  if (H5Dwrite(dset.id, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0)
    return fail("cannot write dataset data");
  return write_params(file.id, params, nparams);
}

int cruna_io_load_restart(const char *fname, int *ns, int *nt) {
  if (!fname || !ns || !nt) return fail("null argument");
  quiet();
  H5Handle file(H5Fopen(fname, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (file.id < 0) return fail(std::string("cannot open ") + fname);
  H5Handle dset(H5Dopen2(file.id, "data", H5P_DEFAULT), H5Dclose);
  if (dset.id < 0) return fail("no dataset data");
  H5Handle space(H5Dget_space(dset.id), H5Sclose);
  hsize_t h = 0;
  if (H5Sget_simple_extent_ndims(space.id) != 1 || H5Sget_simple_extent_dims(space.id, &h, nullptr) < 0 ||
      h != 2)
    return fail("restart data is not (ns, nt)");
  int v[2] = {0, 0};
  if (H5Dread(dset.id, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0)
    return fail("cannot read restart data");
  *ns = v[0];
  *nt = v[1];
  return 0;
}

int cruna_io_param_count(const char *fname, int *count) {
  if (!fname || !count) return fail("null argument");
  std::vector<std::string> rows;
  if (read_params(fname, rows)) return -1;
  int n = 0;
  for (auto &r : rows) n += r.empty() ? 0 : 1;
  *count = n;
  return 0;
}

int cruna_io_param_row(const char *fname, int i, char *out, size_t cap) {
  if (!fname || !out) return fail("null argument");
  std::vector<std::string> rows;
  if (read_params(fname, rows)) return -1;
  if (i < 0 || i >= (int)rows.size()) return fail("row out of range");
  if (rows[i].size() + 1 > cap) return fail("row buffer too small");
  std::memcpy(out, rows[i].c_str(), rows[i].size() + 1);
  return 0;
}

}  // extern "C"
