// Diagnostics (src/Diagnostics): device measurement of one disk.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"

// per measurement: 16 quantities; a radial profile row holds r and the 16,
// an equatorial cell r, phi and the 16 (src/Diagnostics/diagnostics_create_destroy.c:13)
#define DIAG_NS 16
#define DIAG_NV 17
#define DIAG_NE 18

struct DiagArgs {
  int disk, nrg, imin, kmin, kmax;
  double dtout;
  double rz[4];          // R_Min, R_Max, Z_Min, Z_Max (sim_MIN / sim_MAX)
  const int *eq_base;    // per layer k: first equatorial position, or -1
  const int *ioff;       // per interior annulus: cells of the interior annuli before it
  double *vec;           // [nrg][DIAG_NV]
  double *mpart;         // [nrg][5] per-annulus mass sums
  double *scal;          // [DIAG_NS]
  double *equat;         // [N_eq][DIAG_NE]
  double *out_mass;      // inside r=1, near mass 0/1 within 0.2, within 0.4, inner wedge, Mdot probe, its r
};

void launch_diag(const DevMesh &m, const DiagArgs &g, hipStream_t st);
