// Checkpoints of one disk (src/IO): the cell rows packed on the GPU and the
// reference's HDF5 layout: dataset T (the time), GravMass [2][4] (r, phi, M,
// omega of each mass, io_buffer.c:43-54) and Data [Ncells][NUM_Q + 3]
// (io_hdf5.c:15-155).
#include <math.h>
#include <string.h>

#include <vector>

#include "disco_internal.h"
#include "engine.h"
#include "h5.h"

namespace {
bool ckpt_scratch(Engine &e, size_t bytes) {
  if (bytes <= e.ckpt_cap) return true;
  if (e.ckpt_mem) hipFree(e.ckpt_mem);
  e.ckpt_mem = nullptr;
  e.ckpt_cap = 0;
  if (!hip_check(hipMalloc(&e.ckpt_mem, bytes), "checkpoint")) return false;
  e.ckpt_cap = bytes;
  return true;
}
}  // namespace

int Engine::checkpoint_pack(int d, double *rows) {
  if (!ok || d < 0 || d >= ndisk) return -1;
  const DiskParams &P = hdp[d];
  const size_t bytes = 8 * (size_t)nzones[d] * (NQ + 3);
  if (!ckpt_scratch(*this, bytes)) return -1;
  launch_ckpt_pack(mesh(), d, P.NR * P.NZ, (double *)ckpt_mem, st);
  hip_check(hipMemcpyAsync(rows, ckpt_mem, bytes, hipMemcpyDeviceToHost, st), "checkpoint");
  return check_error() ? -1 : 0;
}

int Engine::checkpoint_unpack(int d, const double *rows) {
  if (!ok || d < 0 || d >= ndisk) return -1;
  const DiskParams &P = hdp[d];
  const size_t bytes = 8 * (size_t)nzones[d] * (NQ + 3);
  if (!ckpt_scratch(*this, bytes)) return -1;
  hip_check(hipMemcpyAsync(ckpt_mem, rows, bytes, hipMemcpyHostToDevice, st), "checkpoint");
  launch_ckpt_unpack(mesh(), d, P.NR * P.NZ, (const double *)ckpt_mem, d_tiph[cur], st);
  return check_error() ? -1 : 0;
}

int Engine::checkpoint_write(int d, const char *fname) {
  if (!h5_available()) {
    disco_set_error(15, "checkpoints need the HDF5 library (DISCO_HDF5_LIB or libhdf5.so)");
    return -1;
  }
  std::vector<double> rows((size_t)nzones[d] * (NQ + 3));
  if (checkpoint_pack(d, rows.data())) return -1;
  const DiskScalars s = get_scalars(d);
  double T = s.t, gm[2][4];
  for (int p = 0; p < 2; ++p) {
    gm[p][0] = s.gm[p].r;
    gm[p][1] = s.gm[p].phi;
    gm[p][2] = s.gm[p].M;
    gm[p][3] = s.gm[p].omega;
  }
  const H5Dataset ds[3] = {{"T", 1, {1, 1}, &T},
                           {"GravMass", 2, {2, 4}, &gm[0][0]},
                           {"Data", 2, {(unsigned long long)nzones[d], (unsigned long long)(NQ + 3)}, rows.data()}};
  if (!h5_write_file(fname, ds, 3)) {
    disco_set_error(15, std::string("checkpoint: cannot write ") + fname);
    return -1;
  }
  return 0;
}

// io_hdf5_in + io_readbuf (io_hdf5.c:158-265, io_buffer.c:57-81) into a disk
// whose cells are as cell_create left them: the time, the masses' r, phi, M
// and omega (gravMass_set_chkpt, then gravMass_clean_pi), tiph and the
// primitives; dphi stays what it was (the reference's restart keeps
// cell_create's 2 pi / N_p until the first update recomputes it).
int Engine::checkpoint_read(int d, const char *fname) {
  if (!h5_available()) {
    disco_set_error(15, "checkpoints need the HDF5 library (DISCO_HDF5_LIB or libhdf5.so)");
    return -1;
  }
  long long dims[2];
  double T = 0.0, gm[8];
  const long long n = (long long)nzones[d] * (NQ + 3);
  std::vector<double> rows((size_t)n);
  if (h5_read_dataset(fname, "T", &T, 1, dims) != 1 || h5_read_dataset(fname, "GravMass", gm, 8, dims) != 8 ||
      h5_read_dataset(fname, "Data", rows.data(), n, dims) != n || dims[1] != NQ + 3) {
    disco_set_error(15, std::string("checkpoint: ") + fname + " does not hold this disk's T, GravMass and Data");
    return -1;
  }
  if (checkpoint_unpack(d, rows.data())) return -1;
  DiskScalars s = get_scalars(d);
  s.t = T;
  for (int p = 0; p < 2; ++p) {
    s.gm[p].r = gm[4 * p];
    double phi = gm[4 * p + 1];  // gravMass_clean_pi, GravMass_misc.c:9-17
    while (phi > 2. * M_PI) phi -= 2. * M_PI;
    while (phi < 0.0) phi += 2. * M_PI;
    s.gm[p].phi = phi;
    s.gm[p].M = gm[4 * p + 2];
    s.gm[p].omega = gm[4 * p + 3];
  }
  put_scalars(d, s);
  launch_scalar(mesh(), plan, st, SC_MASS);
  return check_error() ? -1 : 0;
}
