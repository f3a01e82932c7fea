// The analytic frame rotation, src/Sim/sim_w_a.c:8-73, on the host.
//
// Compiled by g++ with -ffp-contract=off, like the reference, so that pow is
// folded or called as there: every frame that does not depend on the binary
// separation is tabulated here once per annulus and the device reads the
// reference's bits.  That matters beyond rounding: with A_POWERLAW the
// epicyclic term kappa^2 = (2 Omega / r) (2 rOm_a + r rdrOm_a) of the split
// fictitious rotation is zero up to rounding, and the sign of that rounding
// residue chooses the branch of cell_add_split_fictitious.c:29-39.
#include <math.h>

#include "disco_internal.h"

#define W_A_MILOS_INDEX 8  // w_a_milos_index, src/Headers/header.h:25

double disco_rOm_a(int type, double r, double a) {
  switch (type) {
    case A_FIXED: return 0.0;
    case A_KEPLER: return pow(r, -0.5);
    case A_OMEGA20: return r * 20;
    case A_MILOS:
      return r * (pow(r, W_A_MILOS_INDEX - 1.5) + pow(a, W_A_MILOS_INDEX - 1.5)) /
             (pow(r, W_A_MILOS_INDEX) + pow(a, W_A_MILOS_INDEX));
    case A_POWERLAW: return pow(r, -1.0);
  }
  disco_set_error(5, "unknown w_analytic_type");
  return 0.0;
}
double disco_rdrOm_a(int type, double r, double a) {
  switch (type) {
    case A_FIXED: return 0.0;
    case A_KEPLER: return -1.5 * pow(r, -1.5);
    case A_OMEGA20: return 0.0;
    case A_MILOS: {
      int n = W_A_MILOS_INDEX;
      double Omega_a = (pow(r, n - 1.5) + pow(a, n - 1.5)) / (pow(r, n) + pow(a, n));
      double arn = pow(a / r, n);
      return (-1.5 + n * arn / (1. + arn)) * Omega_a + (1.5 - n) * pow(a / r, n - 1.5) / (1. + arn) * pow(r, -1.5);
    }
    case A_POWERLAW: return -2. * pow(r, -2.0);
  }
  disco_set_error(5, "unknown w_analytic_type");
  return 0.0;
}
double disco_dtOm_a(int type, double r, double a, double tau) {
  if (type == A_MILOS) {
    double a0 = 1.0;
    double dta_o_a = -.25 * pow(a / a0, -4) / tau;
    int n = W_A_MILOS_INDEX;
    double Omega_a = (pow(r, n - 1.5) + pow(a, n - 1.5)) / (pow(r, n) + pow(a, n));
    double arn = pow(a / r, n);
    if (a > 0.0) return dta_o_a * (Omega_a * (-1.5 + n / (1. + arn)) + (1.5 - n) * pow(r, -1.5) / (1. + arn));
  }
  return 0.0;
}

void disco_frame_rows(const Sim *S, double a, std::vector<double> &out) {
  const int NR = (int)S->N_p.size(), NZ = (int)S->z_faces.size() - 1;
  const int type = S->W_ANALYTIC_TYPE;
  out.assign((size_t)4 * NR * NZ, 0.0);
  for (int k = 0; k < NZ; ++k)
    for (int i = 0; i < NR; ++i) {
      const double rm = S->r_faces[i], rp = S->r_faces[i + 1];
      const double r = .5 * (rp + rm);
      double *o = &out[(size_t)4 * (k * NR + i)];
      o[0] = disco_rOm_a(type, r, a);
      o[1] = disco_rdrOm_a(type, r, a);
      o[2] = disco_dtOm_a(type, r, a, S->OrbShrinkTscale);
      o[3] = disco_rdrOm_a(type, rp, a);
    }
}
