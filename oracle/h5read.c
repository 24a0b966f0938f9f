/*
 * h5read: a plain HDF5 reader for the tests of the .cool writer.
 *
 * TEST INFRASTRUCTURE (the checker, never the product): it reads datasets and root attributes of
 * a file by name through the HDF5 C library, independently of the writer's code, so the tests can
 * rebuild the contact matrix from the file and compare it with the band that was written.
 * Every function returns 0 on success and -1 on any HDF5 error.
 */
#include <hdf5.h>
#include <stdint.h>
#include <string.h>

static hid_t open_ro(const char* path) {
  H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
  return H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
}

/* Length of 1-D dataset dset, its type class (H5T_INTEGER 0, H5T_STRING 3, ...) and element size. */
int h5r_info(const char* path, const char* dset, uint64_t* n, int* type_class, uint64_t* size) {
  hid_t f = open_ro(path);
  if (f < 0) return -1;
  int rc = -1;
  hid_t d = H5Dopen2(f, dset, H5P_DEFAULT);
  if (d >= 0) {
    hid_t s = H5Dget_space(d), t = H5Dget_type(d);
    hsize_t dims = 0;
    if (H5Sget_simple_extent_ndims(s) == 1 && H5Sget_simple_extent_dims(s, &dims, NULL) == 1) {
      *n = dims;
      *type_class = (int)H5Tget_class(t);
      *size = H5Tget_size(t);
      rc = 0;
    }
    H5Tclose(t);
    H5Sclose(s);
    H5Dclose(d);
  }
  H5Fclose(f);
  return rc;
}

static int read_all(const char* path, const char* dset, hid_t mtype, void* out) {
  hid_t f = open_ro(path);
  if (f < 0) return -1;
  int rc = -1;
  hid_t d = H5Dopen2(f, dset, H5P_DEFAULT);
  if (d >= 0) {
    rc = H5Dread(d, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) >= 0 ? 0 : -1;
    H5Dclose(d);
  }
  H5Fclose(f);
  return rc;
}

int h5r_i64(const char* path, const char* dset, int64_t* out) {
  return read_all(path, dset, H5T_NATIVE_INT64, out);
}

int h5r_i32(const char* path, const char* dset, int32_t* out) {
  return read_all(path, dset, H5T_NATIVE_INT32, out);
}

/* Fixed-length strings of `width` bytes each (the dataset's own size). */
int h5r_str(const char* path, const char* dset, uint64_t width, char* out) {
  hid_t t = H5Tcopy(H5T_C_S1);
  H5Tset_size(t, width);
  int rc = read_all(path, dset, t, out);
  H5Tclose(t);
  return rc;
}

int h5r_attr_i64(const char* path, const char* name, int64_t* out) {
  hid_t f = open_ro(path);
  if (f < 0) return -1;
  int rc = -1;
  hid_t a = H5Aopen(f, name, H5P_DEFAULT);
  if (a >= 0) {
    hid_t t = H5Aget_type(a);
    if (H5Tget_class(t) == H5T_INTEGER) rc = H5Aread(a, H5T_NATIVE_INT64, out) >= 0 ? 0 : -1;
    H5Tclose(t);
    H5Aclose(a);
  }
  H5Fclose(f);
  return rc;
}

/* A variable-length string attribute, copied into out (cap bytes, NUL-terminated). */
int h5r_attr_str(const char* path, const char* name, char* out, uint64_t cap) {
  hid_t f = open_ro(path);
  if (f < 0) return -1;
  int rc = -1;
  hid_t a = H5Aopen(f, name, H5P_DEFAULT);
  if (a >= 0) {
    hid_t t = H5Aget_type(a);
    if (H5Tget_class(t) == H5T_STRING && H5Tis_variable_str(t) > 0) {
      char* s = NULL;
      hid_t mt = H5Tcopy(H5T_C_S1);
      H5Tset_size(mt, H5T_VARIABLE);
      H5Tset_cset(mt, H5Tget_cset(t));
      if (H5Aread(a, mt, &s) >= 0 && s) {
        strncpy(out, s, cap - 1);
        out[cap - 1] = '\0';
        H5free_memory(s);
        rc = 0;
      }
      H5Tclose(mt);
    }
    H5Tclose(t);
    H5Aclose(a);
  }
  H5Fclose(f);
  return rc;
}
