// HDF5 files through the library loaded at run time (h5.cpp).
#pragma once

struct H5Dataset {
  const char *name;
  int rank;                 // 1 or 2
  unsigned long long dims[2];
  const double *data;
};

// false when no HDF5 library could be loaded
bool h5_available();
// a new file (truncated) holding the datasets
bool h5_write_file(const char *fname, const H5Dataset *ds, int n);
// dataset `name` of a file: its element count (dims set), -1 on failure;
// out may be null to ask for the size
long long h5_read_dataset(const char *fname, const char *name, double *out, long long max, long long dims[2]);
