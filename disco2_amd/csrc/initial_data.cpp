// The reference's initial-data plugins, src/Cell/initial_data/*.c, chosen by
// InitialDataType (src/Cell/cell_init_function_ptr.c:6-46), and the rows of
// the fixed r boundaries and the boundary damping.
//
// They run once, at set-up, on the host.  This file is compiled by g++ (the
// reference's compiler family) with -ffp-contract=off and calls the same C
// library, so every expression below, kept in the reference's operation
// order, gives the reference's bits: the library folds and calls pow, exp,
// sqrt, cos and sin as the reference build does.
#include "initial_data.h"

#include <math.h>
#include <string.h>

#include <string>

#include "../../include/disco.h"

namespace {

// annulus centre, 0.5 * (sim_FacePos(i-1) + sim_FacePos(i)) (sim_access.c:26-36)
double r_centre(const Sim *S, int i) { return 0.5 * (S->r_faces[i] + S->r_faces[i + 1]); }

double rOm_a(const Sim *S, double r, double a) { return disco_rOm_a(S->W_ANALYTIC_TYPE, r, a); }

// gravMass_dist, src/GravMass/GravMass_access.c:25-31
double grav_dist(const GravMass &g, double r, double phi, double z) {
  double xbh = g.r * cos(g.phi);
  double ybh = g.r * sin(g.phi);
  double x = r * cos(phi);
  double y = r * sin(phi);
  return sqrt((x - xbh) * (x - xbh) + (y - ybh) * (y - ybh) + z * z);
}

void put5(double *p, long long nz, long long c, double a, double b, double v, double w, double u) {
  p[RHO * nz + c] = a;
  p[PPP * nz + c] = b;
  p[URR * nz + c] = v;
  p[UPP * nz + c] = w;
  p[UZZ * nz + c] = u;
}

// ---------------------------------------------------------------- shear
// cell_init_shear.c:57-117 (the single-cell function, :11-55, is the same)
struct Shear {
  double rho = 1.0, Pp = 0.01, v0 = 1.0, t0 = 0.25, nu;
  explicit Shear(const Sim *S) { nu = S->EXPLICIT_VISCOSITY > 0.0 ? S->EXPLICIT_VISCOSITY : 0.05; }
  void cell(const Sim *S, double r, double tiph, double dphi, double *o) const {
    double t = tiph - .5 * dphi;
    double x = r * cos(t) - 3.;
    double y = r * sin(t);
    double vy = 0.0;
    if (fabs(y) < 50.) vy = v0 / sqrt(2. * M_PI * nu * t0) * exp(-x * x / (4. * nu * t0));
    double vr = vy * sin(t);
    double omega = vy * cos(t) / r;
    o[RHO] = rho;
    o[PPP] = Pp;
    o[URR] = vr;
    o[UPP] = omega - rOm_a(S, r, 1.) / r;
    o[UZZ] = 0.0;
  }
};

// ---------------------------------------------------------------- vortex
// cell_init_vortex.c:156-191 (single cell :128-154): cells with r < R only
struct Vortex {
  double rho = 1.0, P0 = 1.0, R = 1.0, A = 0.5, g = 3.0;
  // false where the plugin writes nothing (r >= R)
  bool annulus(double r, double *o) const {
    if (r >= R) return false;
    double Pp = P0 * (1. - A * exp(g * (r / R - R / (R - r) + 1.)));
    double vp = sqrt(A * g * P0 / rho * (2. * R - r) / R * r * r / pow(R - r, 2.) * exp(g * (r / R - R / (R - r) + 1.)));
    o[RHO] = rho;
    o[PPP] = Pp;
    o[URR] = 0.0;
    o[UPP] = vp / r;
    o[UZZ] = 0.0;
    return true;
  }
  static double passive(double tiph, double dphi) {
    double t = tiph - .5 * dphi;
    return cos(t) < 0.0 ? 0.0 : 1.0;
  }
};

// ---------------------------------------------------------------- SStest / middle
// cell_init_SStest.c:49-92 and cell_init_middle.c:49-92 differ in r0, d, the
// floors, v_r and omega; the single-cell functions (:12-47) drop the
// "r < 1: v_r = 0" of the global ones
struct SSLike {
  bool middle;
  double r0;
  int d;
  double rho_min, P_min;
  explicit SSLike(bool mid)
      : middle(mid), r0(mid ? 2.5 : 5.), d(mid ? 10 : 4), rho_min(mid ? 1.e-3 : 1.e-4), P_min(mid ? 1.e-6 : 1.e-7) {}
  void annulus(const Sim *S, double r, bool global, double *o) const {
    double alpha = S->EXPLICIT_VISCOSITY;
    double rho = pow(r, -3. / 5.) * exp(-pow(r / r0, (double)-d));
    double P = S->PoRho_r1 * pow(r, -3. / 2.) * exp(-pow(r / r0, (double)-d));
    double r_drP_o_P = -1.5 + d * pow(r / r0, (double)-d);
    if (rho < rho_min) rho = rho_min;
    if (P < P_min) P = P_min;
    double o2 = 1. / r / r / r * pow(1. + 3. / 16. / r / r, 2.) + r_drP_o_P * P / rho / r / r;
    double omega = sqrt(o2);
    double a = 1.0;
    const int n = 8;  // w_a_milos_index, src/Headers/header.h:25
    omega = (pow(r, (double)n) * omega + pow(a, n - 1.5)) / (pow(r, (double)n) + pow(a, (double)n));
    double vr = middle ? -3. * alpha / r / omega * P / rho * (2. + r_drP_o_P) : 0.0;
    if (global && r < 1.) vr = 0.0;
    o[RHO] = rho;
    o[PPP] = P;
    o[URR] = vr;
    o[UPP] = (middle ? omega : pow(r, -1.5)) - rOm_a(S, r, a) / r;
    o[UZZ] = 0.0;
  }
};

// ---------------------------------------------------------------- rad_dom
// cell_init_rad_dom.c:133-167 (single cell :103-131, the same values)
void rad_dom(const Sim *S, double r, double *o) {
  const double blah = 0.0;  // the file-scope `blah` of :101
  double xi = 0.005;
  double a_o_M = 100;
  double alpha = S->EXPLICIT_VISCOSITY;
  double Gam = S->GAMMALAW;
  double rho0 = 2. / (9. * xi * alpha * Gam) * sqrt(a_o_M);
  double P0 = 8. * xi / (9. * alpha * Gam) * sqrt(a_o_M);
  double rho = rho0 * pow(r, 1.5);
  double P = P0 * pow(r, -1.5);
  double o2 = 1. / r / r / r - 3. / 2. * P / rho / r / r;
  double omega = sqrt(o2);
  double vr = -1.5 * alpha * Gam * (P / rho) * sqrt(r) - blah * pow(r, -4.5);
  o[RHO] = rho;
  o[PPP] = P;
  o[URR] = vr;
  o[UPP] = omega - sqrt(1.) / pow(r, 1.5);
  o[UZZ] = 0.0;
}

// ---------------------------------------------------------------- milos_macfadyen
// single cell, cell_init_milos_macfadyen.c:11-35 (no floor, no potential term)
void milos_single(const Sim *S, double r, double *o) {
  double Mach = 1. / sqrt(S->PoRho_r1);
  double Gam = S->GAMMALAW;
  double rho0 = 1.0;
  double rs = 10.;
  double delta_exp = 3.0;
  double xi_exp = 2.0;
  double rho = rho0 * (pow(rs / r, delta_exp)) * exp(-pow(rs / r, xi_exp));
  double PoRho = 1. / Mach / Mach / Gam;
  o[RHO] = rho;
  o[PPP] = PoRho * rho;
  o[URR] = 0.0;
  o[UPP] = 0.0;
  o[UZZ] = 0.0;
}

// global, cell_init_milos_macfadyen.c:37-152
void milos_global(const Sim *S, const GravMass *gm, const double *tiph, const double *dphi, double *prim,
                  long long nz) {
  double rs = 10.0;
  double rho0 = 1.0;
  double Mach = 1. / sqrt(S->PoRho_r1);
  double delta_exp = 3.0;
  double xi_exp = 2.0;
  double Gam = S->GAMMALAW;
  double cp = 1. / Mach;
  double fac = 0.0001;
  double Mtotal = 1.0;
  double sep = 1.0;
  double massratio = S->MassRatio;
  double M0, M1, r0, r1;
  if (fabs(massratio) > 1.e-8) {
    M0 = Mtotal / (1. + massratio);
    M1 = Mtotal / (1. + 1. / massratio);
    r0 = M1 / Mtotal * sep;
    r1 = M0 / Mtotal * sep;
  } else {
    M0 = 1.0;
    M1 = 0.0;
    r0 = 0.0;
    r1 = 0.0;
  }
  double eps0 = S->G_EPS * r0;
  double eps1 = S->G_EPS * r1;
  double DISK_ALPHA = S->EXPLICIT_VISCOSITY;
  const int NR = (int)S->N_p.size(), NZ = (int)S->z_faces.size() - 1;
  long long c = 0;
  for (int k = 0; k < NZ; ++k)
    for (int i = 0; i < NR; ++i) {
      double r = r_centre(S, i);
      double cs;
      if (r < 1.)
        cs = 1. / Mach;
      else
        cs = sqrt(1. / r) / Mach;
      for (int j = 0; j < S->N_p[i]; ++j, ++c) {
        double phi = tiph[c] - 0.5 * dphi[c];
        double dist_bh0 = grav_dist(gm[0], r, phi, 0.);
        double dist_bh1 = grav_dist(gm[1], r, phi, 0.);
        double cs0, cs1;
        if (dist_bh0 < 0.5)
          cs0 = sqrt(1. / .5) / Mach * sqrt(M0);
        else
          cs0 = sqrt(1. / dist_bh0) / Mach * sqrt(M0);
        if (dist_bh1 < 0.5)
          cs1 = sqrt(1. / .5) / Mach * sqrt(M1);
        else
          cs1 = sqrt(1. / dist_bh1) / Mach * sqrt(M1);
        double PoRho = (cs0 * cs0 + cs1 * cs1) / Gam;
        int n = S->PHI_ORDER;
        double Pot1 = M0 / pow(pow(dist_bh0, (double)n) + pow(eps0, (double)n), 1. / n);
        double Pot2 = M1 / pow(pow(dist_bh1, (double)n) + pow(eps1, (double)n), 1. / n);
        double rho = rho0 * (pow(rs / r, delta_exp)) * exp(-pow(rs / r, xi_exp));
        double omega = sqrt(1.) / pow(r, 1.5);
        omega *= 1. + 3. / 16. / r / r;
        double O2 = omega * omega + cs * cs / r / r * (2. * rs * rs / r / r - 3.);
        double vr;
        if (r < 3.0) {
          omega = 1. / pow(r, 1.5);
          vr = 0.0;
        } else {
          omega = sqrt(O2);
          vr = -3.0 / sqrt(r) * DISK_ALPHA * (1.0 / Mach) * (1.0 / Mach) *
               (1.0 - delta_exp + xi_exp * pow(r / rs, -xi_exp));
        }
        if (rho < (100. * S->RHO_FLOOR)) rho = 100. * S->RHO_FLOOR;
        double Pp = PoRho * rho;
        put5(prim, nz, c, rho * exp(fac * (Pot1 + Pot2) / cp / cp), Pp * exp(fac * (Pot1 + Pot2) / cp / cp), vr,
             omega - sqrt(1.) / pow(r, 1.5), 0.0);
      }
    }
}

// ---------------------------------------------------------------- torus
// cell_init_torus.c:17-90: srand(666 + rank), one draw per cell
void torus_global(const Sim *S, double *prim, long long nz) {
  double GAMMALAW = S->GAMMALAW;
  double M = 1.0;
  double r_inner = 0.7;
  double r_max = 1.0;
  double q = 2.;
  double s = 2. * q - q;
  double GammaFac = (GAMMALAW - 1.) / GAMMALAW;
  double Phi_inner = -1.0 / r_inner;
  double Phi_max = -1.0 / r_max;
  double Rho_norm =
      pow(GammaFac * (Phi_inner - Phi_max + pow(r_inner, -s) / s - pow(1.0, -s) / s), 1.0 / (GAMMALAW - 1.0));
  const double RAND_MAX_D = 2147483647.0;
  GlibcRand g(666);
  const int NR = (int)S->N_p.size(), NZ = (int)S->z_faces.size() - 1;
  long long c = 0;
  for (int k = 0; k < NZ; ++k) {
    double z;
    if (S->N_global[1] != 1) {
      double zp = S->z_faces[k + 1];
      double zm = S->z_faces[k];
      z = 0.5 * (zm + zp);
    } else {
      z = 0.0;
    }
    for (int i = 0; i < NR; ++i) {
      double r = r_centre(S, i);
      double R = sqrt(r * r + z * z);
      double Phi = -M / R;
      double omega = pow(r, -q);
      double PoRho = GammaFac * (Phi_inner - Phi + pow(r_inner, -s) / s - pow(r, -s) / s);
      if (PoRho < 1.0e-5) PoRho = 1.0e-5;
      double rho = pow(PoRho, 1.0 / (GAMMALAW - 1.0)) / Rho_norm;
      double Pp = pow(PoRho, 1.0 / GammaFac) / Rho_norm;
      if (PoRho < 1.0e-5) Pp = rho;
      for (int j = 0; j < S->N_p[i]; ++j, ++c) {
        double delta = .02 * ((double)g.next() / RAND_MAX_D - .5);
        put5(prim, nz, c, rho, Pp, 0.0, omega * (1. + delta) - pow(r, -2.), 0.0);
      }
    }
  }
}

bool valid_type(int type, const char *what) {
  if (type < ID_SHEAR || type > ID_MIDDLE) {
    disco_set_error(8, std::string("not a valid ") + what + " initial data routine choice");
    return false;
  }
  return true;
}

}  // namespace

bool disco_cell_init(const Sim *S, int type, const GravMass *g, const double *tiph, const double *dphi, double *prim,
                     long long nz, int NQ) {
  if (!valid_type(type, "global")) return false;
  GravMass gm[2];
  if (g) {
    gm[0] = g[0];
    gm[1] = g[1];
  } else {  // src/main.c:103: the masses are set before the cells
    memset(gm, 0, sizeof(gm));
    (*gravMass_init_ptr(const_cast<Sim *>(S)))(gm, const_cast<Sim *>(S));
  }
  const int NR = (int)S->N_p.size(), NZ = (int)S->z_faces.size() - 1;
  double o[5];
  switch (type) {
    case ID_MILOS_MACFADYEN: milos_global(S, gm, tiph, dphi, prim, nz); return true;
    case ID_TORUS: torus_global(S, prim, nz); return true;
    default: break;
  }
  const Shear shear(S);
  const Vortex vortex;
  const SSLike ss(type == ID_MIDDLE);
  long long c = 0;
  for (int k = 0; k < NZ; ++k)
    for (int i = 0; i < NR; ++i) {
      const double r = r_centre(S, i);
      bool set = true;
      if (type == ID_VORTEX)
        set = vortex.annulus(r, o);
      else if (type == ID_SSTEST || type == ID_MIDDLE)
        ss.annulus(S, r, true, o);
      else if (type == ID_RAD_DOM)
        rad_dom(S, r, o);
      for (int j = 0; j < S->N_p[i]; ++j, ++c) {
        if (type == ID_SHEAR) shear.cell(S, r, tiph[c], dphi[c], o);
        if (!set) continue;
        put5(prim, nz, c, o[RHO], o[PPP], o[URR], o[UPP], o[UZZ]);
        if (type == ID_VORTEX && S->NUM_C < NQ) prim[S->NUM_C * nz + c] = Vortex::passive(tiph[c], dphi[c]);
      }
    }
  return true;
}

int disco_single_init(const Sim *S, int type, int i, double tiph, double dphi, double *prim, int NQ) {
  if (!valid_type(type, "single point")) return -1;
  const double r = r_centre(S, i);
  switch (type) {
    case ID_SHEAR: Shear(S).cell(S, r, tiph, dphi, prim); return 5;
    case ID_VORTEX:
      if (!Vortex().annulus(r, prim)) return 0;
      if (S->NUM_C < NQ) prim[S->NUM_C] = Vortex::passive(tiph, dphi);
      return 5;
    case ID_SSTEST:
    case ID_MIDDLE: SSLike(S->InitialDataType == ID_MIDDLE).annulus(S, r, false, prim); return 5;
    case ID_MILOS_MACFADYEN: milos_single(S, r, prim); return 5;
    case ID_RAD_DOM: rad_dom(S, r, prim); return 5;
  }
  // cell_init_torus.c:12-15 prints a warning and exits
  disco_set_error(8, "the torus initial data has no single-cell function (the reference exits)");
  return -1;
}

bool disco_init_rows(const Sim *S, int NQ, std::vector<InitRow> &rows, bool *any, bool values_always) {
  const int NR = (int)S->N_p.size(), NZ = (int)S->z_faces.size() - 1;
  const bool fixed_r = S->BoundTypeR == BOUND_FIXED;
  const bool fixed_z = S->N_global[1] > 1 && S->BoundTypeZ == BOUND_FIXED;
  const bool damp = S->DAMP_TIME > 0.0;
  rows.assign((size_t)NR * NZ, InitRow());
  for (auto &w : rows) memset(&w, 0, sizeof(w));
  bool any_ = false;
  if (!any) any = &any_;
  *any = false;
  if (!fixed_r && !fixed_z && !damp && !values_always) return true;
  if (!valid_type(S->InitialDataType, "single point")) return false;
  const int type = S->InitialDataType;
  if (type == ID_TORUS) {
    disco_set_error(8, "fixed boundaries and damping call the single-cell initial data, which the torus lacks "
                       "(the reference exits)");
    return false;
  }
  if (damp && NQ != 5) {
    disco_set_error(8, "damping with passive scalars reads initial values the reference never sets");
    return false;
  }
  if (damp && type == ID_SHEAR) {
    disco_set_error(8, "damping with the shear initial data reads the angle of a cell the reference never sets");
    return false;
  }
  // cell_bc_damp, src/Cell/cell_misc.c:98-132
  const double RMIN = S->MIN[0], RMAX = S->MAX[0];
  const double R0 = 1. / S->DAMP_TIME / 2. / M_PI;
  const Shear shear(S);
  for (int i = 0; i < NR; ++i) {
    InitRow w;
    memset(&w, 0, sizeof(w));
    const double r = r_centre(S, i);
    if (fixed_r && ((i < S->Nghost_min[0] && S->NoInnerBC != 1) || i >= NR - S->Nghost_max[0])) w.flags |= INIT_FIXED;
    if (damp) {
      double rp = S->r_faces[i + 1], rm = S->r_faces[i];
      double rr = .5 * (rp + rm);
      if (rr < S->RDAMP_INNER || rr > S->RDAMP_OUTER) {
        double rate = R0 * (rr - S->RDAMP_INNER) / (RMIN - S->RDAMP_INNER);
        if (rr > S->RDAMP_OUTER) rate = R0 * (rr - S->RDAMP_OUTER) / (RMAX - S->RDAMP_OUTER);
        w.rate = rate;
        w.flags |= INIT_DAMP;
      }
    }
    if (type == ID_SHEAR) {
      w.a = rOm_a(S, r, 1.) / r;
      w.c0 = shear.v0 / sqrt(2. * M_PI * shear.nu * shear.t0);
      w.c1 = 4. * shear.nu * shear.t0;
      w.flags |= INIT_SET;
    } else {
      double o[DISCO_MAXQ];
      int n = disco_single_init(S, type, i, 0.0, 0.0, o, 5);
      if (n < 0) return false;
      if (n == 5) {
        for (int q = 0; q < 5; ++q) w.prim[q] = o[q];
        w.flags |= INIT_SET;
      }
    }
    if ((w.flags & INIT_DAMP) && !(w.flags & INIT_SET)) {
      disco_set_error(8, "damping where the vortex initial data sets nothing (r >= 1) reads values the reference "
                         "never sets");
      return false;
    }
    for (int k = 0; k < NZ; ++k) rows[(size_t)k * NR + i] = w;
    if (w.flags & (INIT_FIXED | INIT_DAMP)) *any = true;
  }
  return true;
}
