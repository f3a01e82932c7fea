// Host side of the diagnostics: one device measurement of a disk
// (Engine::diag_measure), and the equatorial HDF5 dump.
// The class API (diagnostics_create / _set / _print / _destroy,
// src/Headers/Diagnostics.h:29-38) is in capi.cpp on top of these.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "diagnostics.h"
#include "disco_internal.h"
#include "engine.h"
#include "h5.h"

int Engine::diag_sizes(int d, int *nrg, int *neq) const {
  const DiskParams &P = hdp[d];
  const int imin = P.NgMinR, imax = P.NR - P.NgMaxR, kmin = P.NgMinZ, kmax = P.NZ - P.NgMaxZ;
  *nrg = imax - imin;
  int n = 0;  // diagnostics_create_destroy.c:34-51
  for (int k = kmin; k < kmax; ++k) {
    const AnnInfo &A = hann[P.ann_base + k * P.NR];
    const double zp = A.zp, z = 0.5 * (A.zm + A.zp);
    if ((fabs(zp) < 0.0000001) || (fabs(z) < 0.0000001))
      for (int i = imin; i < imax; ++i) n += hann[P.ann_base + k * P.NR + i].N;
  }
  *neq = n;
  return 0;
}

int Engine::diag_measure(int d, double dtout, const double *rz, double *vec, double *scal, double *equat,
                         double *mdot) {
  if (!ok) return -1;
  if (d < 0 || d >= ndisk) {
    disco_set_error(14, "diagnostics: no such disk");
    return -1;
  }
  const DiskParams &P = hdp[d];
  if (P.cooling == 1) {
    disco_set_error(14, "diagnostics: the Cool column needs Cooling 0 (the engine keeps no per-cell Cool)");
    return -1;
  }
  int nrg, neq;
  diag_sizes(d, &nrg, &neq);
  const int imin = P.NgMinR, kmin = P.NgMinZ, kmax = P.NZ - P.NgMaxZ;
  std::vector<int> eqb(P.NZ, -1), ioff(nrg, 0);
  {
    int pos = 0, per_layer = 0;
    for (int ii = 0; ii < nrg; ++ii) {
      ioff[ii] = per_layer;
      per_layer += hann[P.ann_base + kmin * P.NR + imin + ii].N;
    }
    for (int k = kmin; k < kmax; ++k) {
      const AnnInfo &A = hann[P.ann_base + k * P.NR];
      const double zp = A.zp, z = 0.5 * (A.zm + A.zp);
      if ((fabs(zp) < 0.0000001) || (fabs(z) < 0.0000001)) {
        eqb[k] = pos;
        pos += per_layer;
      }
    }
  }
  // device scratch: layer bases, annulus offsets, profiles, mass partials,
  // scalars, equatorial cells, masses
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o_eqb = 0, o_ioff = o_eqb + al(sizeof(int) * P.NZ), o_vec = o_ioff + al(sizeof(int) * nrg),
               o_mp = o_vec + al(8 * (size_t)nrg * DIAG_NV), o_sc = o_mp + al(8 * (size_t)nrg * 5),
               o_eq = o_sc + al(8 * DIAG_NS), o_out = o_eq + al(8 * (size_t)neq * DIAG_NE), total = o_out + al(8 * 8);
  if (total > diag_cap) {
    if (diag_mem) hipFree(diag_mem);
    diag_mem = nullptr;
    diag_cap = 0;
    if (!hip_check(hipMalloc(&diag_mem, total), "diagnostics")) return -1;
    diag_cap = total;
  }
  char *b = (char *)diag_mem;
  DiagArgs g;
  g.disk = d;
  g.nrg = nrg;
  g.imin = imin;
  g.kmin = kmin;
  g.kmax = kmax;
  g.dtout = dtout;
  for (int x = 0; x < 4; ++x) g.rz[x] = rz[x];
  g.eq_base = (const int *)(b + o_eqb);
  g.ioff = (const int *)(b + o_ioff);
  g.vec = (double *)(b + o_vec);
  g.mpart = (double *)(b + o_mp);
  g.scal = (double *)(b + o_sc);
  g.equat = (double *)(b + o_eq);
  g.out_mass = (double *)(b + o_out);
  hip_check(hipMemcpyAsync(b + o_eqb, eqb.data(), sizeof(int) * P.NZ, hipMemcpyHostToDevice, st), "diagnostics");
  if (nrg) hip_check(hipMemcpyAsync(b + o_ioff, ioff.data(), sizeof(int) * nrg, hipMemcpyHostToDevice, st), "diagnostics");
  launch_diag(mesh(), g, st);
  double out[8];
  hip_check(hipMemcpyAsync(vec, g.vec, 8 * (size_t)nrg * DIAG_NV, hipMemcpyDeviceToHost, st), "diagnostics");
  hip_check(hipMemcpyAsync(scal, g.scal, 8 * DIAG_NS, hipMemcpyDeviceToHost, st), "diagnostics");
  if (neq && equat)
    hip_check(hipMemcpyAsync(equat, g.equat, 8 * (size_t)neq * DIAG_NE, hipMemcpyDeviceToHost, st), "diagnostics");
  diag_equat_dev = g.equat;
  diag_equat_n = (size_t)neq * DIAG_NE;
  hip_check(hipMemcpyAsync(out, g.out_mass, 8 * 8, hipMemcpyDeviceToHost, st), "diagnostics");
  if (check_error()) return -1;
  const DiskScalars s = get_scalars(d);
  mdot[0] = out[6];  // Mdot_near_req1 (the reference's probe reads the v_z profile column)
  mdot[1] = out[7];  // r_near_req1
  mdot[2] = out[0];  // mass inside r = 1
  mdot[3] = out[5];  // inner wedge
  mdot[4] = s.gm[0].Mdot;
  mdot[5] = s.gm[1].Mdot;
  mdot[6] = s.gm[0].Macc;
  mdot[7] = s.gm[1].Macc;
  mdot[8] = out[1];
  mdot[9] = out[2];
  mdot[10] = out[3];
  mdot[11] = out[4];
  return 0;
}

// the equatorial cells of the last measurement (they stay on the device
// until asked for: the reference writes them only when it dumps)
int Engine::diag_equat(double *out) {
  if (!diag_equat_dev) {
    disco_set_error(14, "diagnostics: nothing measured yet");
    return -1;
  }
  if (diag_equat_n)
    hip_check(hipMemcpyAsync(out, diag_equat_dev, 8 * diag_equat_n, hipMemcpyDeviceToHost, st), "diagnostics");
  return check_error() ? -1 : 0;
}

// The equatorial dump of diagnostics_print (diagnostics_print.c:40-80):
// dataset EQUAT of n0 x n1 doubles
bool diag_h5_write(const char *fname, const double *data, size_t n0, size_t n1) {
  const H5Dataset ds = {"EQUAT", 2, {(unsigned long long)n0, (unsigned long long)n1}, data};
  return h5_write_file(fname, &ds, 1);
}
