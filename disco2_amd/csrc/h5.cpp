// HDF5 files of the diagnostics and the checkpoints, through the HDF5
// library loaded at run time (DISCO_HDF5_LIB, else libhdf5.so on the loader
// path or in /opt/conda/lib): the library has no link-time dependency on it.
// Datasets are plain (contiguous, H5P_DEFAULT) 1- or 2-D arrays of doubles,
// the form the reference's serial and collective writes leave in the file.
#include "h5.h"

#include <dlfcn.h>
#include <stdint.h>
#include <stdlib.h>

namespace {

typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;

struct H5Api {
  bool tried = false, ok = false;
  herr_t (*open)(void);
  hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
  hid_t (*fopen)(const char *, unsigned, hid_t);
  hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
  hid_t (*dcreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
  hid_t (*dopen2)(hid_t, const char *, hid_t);
  hid_t (*dget_space)(hid_t);
  int (*sget_ndims)(hid_t);
  int (*sget_dims)(hid_t, hsize_t *, hsize_t *);
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
  herr_t (*dclose)(hid_t);
  herr_t (*sclose)(hid_t);
  herr_t (*fclose)(hid_t);
  hid_t *native_double;
};

template <class T>
void sym(void *L, T &f, const char *name) {
  f = (T)dlsym(L, name);
}

H5Api &h5() {
  static H5Api a;
  if (a.tried) return a;
  a.tried = true;
  const char *env = getenv("DISCO_HDF5_LIB");
  const char *names[] = {env, "libhdf5.so.103", "libhdf5.so", "/opt/conda/lib/libhdf5.so.103",
                         "/opt/conda/lib/libhdf5.so"};
  void *L = nullptr;
  for (const char *n : names)
    if (n && (L = dlopen(n, RTLD_NOW | RTLD_LOCAL))) break;
  if (!L) return a;
  sym(L, a.open, "H5open");
  sym(L, a.fcreate, "H5Fcreate");
  sym(L, a.fopen, "H5Fopen");
  sym(L, a.screate_simple, "H5Screate_simple");
  sym(L, a.dcreate2, "H5Dcreate2");
  sym(L, a.dopen2, "H5Dopen2");
  sym(L, a.dget_space, "H5Dget_space");
  sym(L, a.sget_ndims, "H5Sget_simple_extent_ndims");
  sym(L, a.sget_dims, "H5Sget_simple_extent_dims");
  sym(L, a.dwrite, "H5Dwrite");
  sym(L, a.dread, "H5Dread");
  sym(L, a.dclose, "H5Dclose");
  sym(L, a.sclose, "H5Sclose");
  sym(L, a.fclose, "H5Fclose");
  a.native_double = (hid_t *)dlsym(L, "H5T_NATIVE_DOUBLE_g");
  a.ok = a.open && a.fcreate && a.fopen && a.screate_simple && a.dcreate2 && a.dopen2 && a.dget_space &&
         a.sget_ndims && a.sget_dims && a.dwrite && a.dread && a.dclose && a.sclose && a.fclose && a.native_double &&
         a.open() >= 0;
  return a;
}

const hid_t P_DEFAULT = 0, S_ALL = 0;
const unsigned ACC_RDONLY = 0x0000u, ACC_TRUNC = 0x0002u;

}  // namespace

bool h5_available() { return h5().ok; }

bool h5_write_file(const char *fname, const H5Dataset *ds, int n) {
  H5Api &a = h5();
  if (!a.ok) return false;
  hid_t f = a.fcreate(fname, ACC_TRUNC, P_DEFAULT, P_DEFAULT);
  if (f < 0) return false;
  bool good = true;
  for (int k = 0; k < n && good; ++k) {
    hsize_t dims[2] = {(hsize_t)ds[k].dims[0], (hsize_t)ds[k].dims[1]};
    hid_t fs = a.screate_simple(ds[k].rank, dims, nullptr), ms = a.screate_simple(ds[k].rank, dims, nullptr);
    hid_t d = a.dcreate2(f, ds[k].name, *a.native_double, fs, P_DEFAULT, P_DEFAULT, P_DEFAULT);
    a.sclose(fs);
    good = d >= 0;
    if (good) {
      fs = a.dget_space(d);
      good = a.dwrite(d, *a.native_double, ms, fs, P_DEFAULT, ds[k].data) >= 0;
      a.sclose(fs);
      a.dclose(d);
    }
    a.sclose(ms);
  }
  a.fclose(f);
  return good;
}

long long h5_read_dataset(const char *fname, const char *name, double *out, long long max, long long dims[2]) {
  H5Api &a = h5();
  if (!a.ok) return -1;
  hid_t f = a.fopen(fname, ACC_RDONLY, P_DEFAULT);
  if (f < 0) return -1;
  long long n = -1;
  hid_t d = a.dopen2(f, name, P_DEFAULT);
  if (d >= 0) {
    hid_t fs = a.dget_space(d);
    hsize_t hd[2] = {1, 1};
    int rank = a.sget_ndims(fs);
    if (rank >= 1 && rank <= 2 && a.sget_dims(fs, hd, nullptr) >= 0) {
      if (rank == 1) hd[1] = 1;
      dims[0] = (long long)hd[0];
      dims[1] = (long long)hd[1];
      const long long cnt = dims[0] * dims[1];
      if (out && cnt <= max && a.dread(d, *a.native_double, S_ALL, S_ALL, P_DEFAULT, out) >= 0)
        n = cnt;
      else if (!out)
        n = cnt;
    }
    a.sclose(fs);
    a.dclose(d);
  }
  a.fclose(f);
  return n;
}
