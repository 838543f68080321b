// Snapshot and field-output files over the HDF5 C library (include/coffee_h5.h).
//
// Serial HDF5: one GPU, one process, so the reference's MPI-IO file access
// (hdf_wrapper.cpp:27-45, 72-82) reduces to the plain file driver; the file
// layout (dataset names, reversed dimensions, element types) is the same.
#include "coffee_h5.h"

#include <hdf5.h>

#include <cstring>
#include <string>

struct coffee_h5_file {
  hid_t id = -1;
};

namespace {

std::string& last_error() {
  static thread_local std::string e;
  return e;
}

int fail(int code, const std::string& msg) {
  last_error() = msg;
  return code;
}

// HDF5 prints its error stack to stderr by default; the ABI reports errors
// through return codes and coffee_h5_last_error instead.
void quiet() {
  static bool done = false;
  if (!done) {
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    done = true;
  }
}

hid_t native(int type) {
  switch (type) {
    case COFFEE_H5_F32: return H5T_NATIVE_FLOAT;
    case COFFEE_H5_F64: return H5T_NATIVE_DOUBLE;
    case COFFEE_H5_U32: return H5T_NATIVE_UINT32;
    case COFFEE_H5_I32: return H5T_NATIVE_INT32;
    default: return -1;
  }
}

// Element class and size of the dataset match the requested type.
bool same_type(hid_t dtype, int type) {
  const H5T_class_t cls = H5Tget_class(dtype);
  const size_t sz = H5Tget_size(dtype);
  switch (type) {
    case COFFEE_H5_F32: return cls == H5T_FLOAT && sz == 4;
    case COFFEE_H5_F64: return cls == H5T_FLOAT && sz == 8;
    case COFFEE_H5_U32: return cls == H5T_INTEGER && sz == 4 && H5Tget_sign(dtype) == H5T_SGN_NONE;
    case COFFEE_H5_I32: return cls == H5T_INTEGER && sz == 4 && H5Tget_sign(dtype) == H5T_SGN_2;
    default: return false;
  }
}

int check_args(coffee_h5_file* f, const char* name, int type) {
  if (!f || f->id < 0 || !name || !*name) return fail(COFFEE_H5_ERR_INVALID, "null file or name");
  if (native(type) < 0) return fail(COFFEE_H5_ERR_INVALID, "unknown element type");
  return COFFEE_H5_OK;
}

// Closes HDF5 identifiers at scope exit.
struct Closer {
  hid_t id;
  herr_t (*fn)(hid_t);
  ~Closer() {
    if (id >= 0) fn(id);
  }
};

}  // namespace

extern "C" {

const char* coffee_h5_last_error(void) { return last_error().c_str(); }

int coffee_h5_create(const char* path, coffee_h5_file** out) {
  if (!path || !out) return fail(COFFEE_H5_ERR_INVALID, "null argument");
  quiet();
  *out = nullptr;
  hid_t id = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (id < 0) return fail(COFFEE_H5_ERR_IO, std::string("cannot create ") + path);
  *out = new coffee_h5_file{id};
  return COFFEE_H5_OK;
}

int coffee_h5_open(const char* path, int writable, coffee_h5_file** out) {
  if (!path || !out) return fail(COFFEE_H5_ERR_INVALID, "null argument");
  quiet();
  *out = nullptr;
  hid_t id = H5Fopen(path, writable ? H5F_ACC_RDWR : H5F_ACC_RDONLY, H5P_DEFAULT);
  if (id < 0) return fail(COFFEE_H5_ERR_IO, std::string("cannot open ") + path);
  *out = new coffee_h5_file{id};
  return COFFEE_H5_OK;
}

int coffee_h5_close(coffee_h5_file* f) {
  if (!f) return COFFEE_H5_OK;
  herr_t e = f->id >= 0 ? H5Fclose(f->id) : 0;
  delete f;
  if (e < 0) return fail(COFFEE_H5_ERR_IO, "H5Fclose failed");
  return COFFEE_H5_OK;
}

int coffee_h5_write_array(coffee_h5_file* f, const char* name, const void* host, int type,
                          int ndim, const unsigned long long ext[3]) {
  int rc = check_args(f, name, type);
  if (rc) return rc;
  if (!host || !ext || ndim < 1 || ndim > 3) return fail(COFFEE_H5_ERR_INVALID, "bad array");
  hsize_t dims[3];
  for (int i = 0; i < ndim; i++) {
    if (ext[i] == 0) return fail(COFFEE_H5_ERR_INVALID, "zero extent");
    dims[i] = (hsize_t)ext[ndim - 1 - i];
  }
  Closer space{H5Screate_simple(ndim, dims, nullptr), H5Sclose};
  if (space.id < 0) return fail(COFFEE_H5_ERR_IO, "H5Screate_simple failed");
  Closer ds{H5Dcreate2(f->id, name, native(type), space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
            H5Dclose};
  if (ds.id < 0) return fail(COFFEE_H5_ERR_IO, std::string("cannot create dataset ") + name);
  if (H5Dwrite(ds.id, native(type), H5S_ALL, H5S_ALL, H5P_DEFAULT, host) < 0)
    return fail(COFFEE_H5_ERR_IO, std::string("cannot write dataset ") + name);
  return COFFEE_H5_OK;
}

int coffee_h5_read_array(coffee_h5_file* f, const char* name, void* host, int type, int ndim,
                         const unsigned long long ext[3]) {
  int rc = check_args(f, name, type);
  if (rc) return rc;
  if (!host || !ext || ndim < 1 || ndim > 3) return fail(COFFEE_H5_ERR_INVALID, "bad array");
  Closer ds{H5Dopen2(f->id, name, H5P_DEFAULT), H5Dclose};
  if (ds.id < 0) return fail(COFFEE_H5_ERR_IO, std::string("no dataset ") + name);
  Closer dt{H5Dget_type(ds.id), H5Tclose};
  if (dt.id < 0 || !same_type(dt.id, type))
    return fail(COFFEE_H5_ERR_SHAPE, std::string("dataset ") + name + ": element type differs");
  Closer space{H5Dget_space(ds.id), H5Sclose};
  if (space.id < 0) return fail(COFFEE_H5_ERR_IO, "H5Dget_space failed");
  const int r = H5Sget_simple_extent_ndims(space.id);
  if (r != ndim) return fail(COFFEE_H5_ERR_SHAPE, std::string("dataset ") + name + ": rank differs");
  hsize_t dims[3] = {0, 0, 0};
  H5Sget_simple_extent_dims(space.id, dims, nullptr);
  for (int i = 0; i < ndim; i++)
    if (dims[i] != (hsize_t)ext[ndim - 1 - i])
      return fail(COFFEE_H5_ERR_SHAPE, std::string("dataset ") + name + ": extent differs");
  if (H5Dread(ds.id, native(type), H5S_ALL, H5S_ALL, H5P_DEFAULT, host) < 0)
    return fail(COFFEE_H5_ERR_IO, std::string("cannot read dataset ") + name);
  return COFFEE_H5_OK;
}

int coffee_h5_write_scalar(coffee_h5_file* f, const char* name, int type, const void* value) {
  int rc = check_args(f, name, type);
  if (rc) return rc;
  if (!value) return fail(COFFEE_H5_ERR_INVALID, "null value");
  Closer space{H5Screate(H5S_SCALAR), H5Sclose};
  if (space.id < 0) return fail(COFFEE_H5_ERR_IO, "H5Screate failed");
  Closer ds{H5Dcreate2(f->id, name, native(type), space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
            H5Dclose};
  if (ds.id < 0) return fail(COFFEE_H5_ERR_IO, std::string("cannot create dataset ") + name);
  if (H5Dwrite(ds.id, native(type), H5S_ALL, H5S_ALL, H5P_DEFAULT, value) < 0)
    return fail(COFFEE_H5_ERR_IO, std::string("cannot write dataset ") + name);
  return COFFEE_H5_OK;
}

int coffee_h5_read_scalar(coffee_h5_file* f, const char* name, int type, void* value) {
  int rc = check_args(f, name, type);
  if (rc) return rc;
  if (!value) return fail(COFFEE_H5_ERR_INVALID, "null value");
  Closer ds{H5Dopen2(f->id, name, H5P_DEFAULT), H5Dclose};
  if (ds.id < 0) return fail(COFFEE_H5_ERR_IO, std::string("no dataset ") + name);
  Closer space{H5Dget_space(ds.id), H5Sclose};
  if (space.id < 0 || H5Sget_simple_extent_npoints(space.id) != 1)
    return fail(COFFEE_H5_ERR_SHAPE, std::string("dataset ") + name + " is not a scalar");
  if (H5Dread(ds.id, native(type), H5S_ALL, H5S_ALL, H5P_DEFAULT, value) < 0)
    return fail(COFFEE_H5_ERR_IO, std::string("cannot read dataset ") + name);
  return COFFEE_H5_OK;
}

int coffee_h5_has(coffee_h5_file* f, const char* name) {
  if (!f || f->id < 0 || !name) return fail(COFFEE_H5_ERR_INVALID, "null file or name");
  htri_t e = H5Lexists(f->id, name, H5P_DEFAULT);
  if (e < 0) return fail(COFFEE_H5_ERR_IO, "H5Lexists failed");
  return e > 0 ? 1 : 0;
}

}  // extern "C"
