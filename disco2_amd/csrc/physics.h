// Per-zone and per-face FP64 physics of the hydro substep, as device
// functions.  Expression association follows the reference exactly (the
// library is compiled with -ffp-contract=off), so a kernel that adds the
// same terms in the same order reproduces the reference bit for bit,
// except where libm transcendentals (sin, cos, pow, exp) differ by an ulp.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "disco_types.h"

// every function here also compiles for the host, where tests check the
// hand-written derivatives against finite differences of these forms
#define DHD __host__ __device__

#define TWO_PI (2.0 * M_PI)  // the reference's PHIMAX, src/Headers/header.h:21
#define HALF_TURN (TWO_PI / 2.)
#define MAXWRAP 64            // angle wraps beyond this mean a diverged state

// while-loop angle wraps of the reference, with a bounded trip count.  The
// loops almost never run more than once, so one trip is taken with selects;
// when its result is in range it is what the loops return (one trip of the
// first loop, or of the second, ending both), and otherwise the loops run
// from the start: a rarely taken branch instead of two divergent loops in
// every call, with the same result.
DHD __forceinline__ double wrap_range(double x, double lo, double hi) {
  double y = x > hi ? x - TWO_PI : (x < lo ? x + TWO_PI : x);
  if (y > hi || y < lo) {
    y = x;
    int n = 0;
#pragma unroll 1
    while (y > hi && n++ < MAXWRAP) y -= TWO_PI;
#pragma unroll 1
    while (y < lo && n++ < MAXWRAP) y += TWO_PI;
  }
  return y;
}
DHD __forceinline__ double clean_pi(double phi) {  // src/Cell/cell_misc.c:19-20
  return wrap_range(phi, 0.0, TWO_PI);
}
DHD __forceinline__ double wrap_pm(double d) {  // e.g. src/Cell/cell_plm.c:39-40
  return wrap_range(d, -HALF_TURN, HALF_TURN);
}
DHD __forceinline__ double wrap_p(double d) {  // e.g. src/Face/face_create_destroy.c:88-89
  return wrap_range(d, 0.0, TWO_PI);
}

// analytic frame rotation, src/Sim/sim_w_a.c:8-73
#define W_A_MILOS_INDEX 8
static DHD __noinline__ double d_rOm_a_any(int type, double r, double a) {
  switch (type) {
    case A_KEPLER: return pow(r, -0.5);
    case A_OMEGA20: return r * 20;
    case A_MILOS:
      return r * (pow(r, W_A_MILOS_INDEX - 1.5) + pow(a, W_A_MILOS_INDEX - 1.5)) /
             (pow(r, W_A_MILOS_INDEX) + pow(a, W_A_MILOS_INDEX));
    case A_POWERLAW: return pow(r, -1.0);
    default: return 0.0;
  }
}
static DHD __noinline__ double d_rdrOm_a_any(int type, double r, double a) {
  switch (type) {
    case A_KEPLER: return -1.5 * pow(r, -1.5);
    case A_MILOS: {
      int n = W_A_MILOS_INDEX;
      double Omega_a = (pow(r, n - 1.5) + pow(a, n - 1.5)) / (pow(r, n) + pow(a, n));
      double arn = pow(a / r, n);
      return (-1.5 + n * arn / (1. + arn)) * Omega_a + (1.5 - n) * pow(a / r, n - 1.5) / (1. + arn) * pow(r, -1.5);
    }
    case A_POWERLAW: return -2. * pow(r, -2.0);
    default: return 0.0;
  }
}
// the common non-rotating frame inline, the rest out of line
DHD __forceinline__ double d_rOm_a(int type, double r, double a) {
  return type == A_FIXED ? 0.0 : d_rOm_a_any(type, r, a);
}
DHD __forceinline__ double d_rdrOm_a(int type, double r, double a) {
  return type == A_FIXED ? 0.0 : d_rdrOm_a_any(type, r, a);
}
DHD __forceinline__ double d_dtOm_a(int type, double r, double a, double tau) {
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

// gravMass_dist, src/GravMass/GravMass_access.c:25-31.  The mass position
// (xbh, ybh) = (r_p cos phi_p, r_p sin phi_p) is uniform over a launch and is
// computed once per disk and substep (DiskScalars::xbh/ybh); cphi, sphi are
// cos and sin of the cell angle, computed once per cell.
DHD __forceinline__ double d_grav_dist(double xbh, double ybh, double r, double cphi, double sphi,
                                             double z) {
  double x = r * cphi;
  double y = r * sphi;
  return sqrt((x - xbh) * (x - xbh) + (y - ybh) * (y - ybh) + z * z);
}

// x^-1.5 and x^1.5 through the correctly rounded sqrt (within two ulps of
// pow); the reference calls pow (src/Riemann/Riemann_routines.c:262,
// src/Cell/cell_source.c:12) and device and host pow differ by ulps anyway.
DHD __forceinline__ double pow_m1p5(double x) { return 1.0 / (x * sqrt(x)); }
DHD __forceinline__ double pow_p1p5(double x) { return x * sqrt(x); }

// src/Cell/cell_conversion.c:11-33
template <int NQ>
DHD __forceinline__ void d_prim2cons(const double *prim, double *cons, double r, double dV,
                                            double G, int NC) {
  double rho = prim[RHO];
  double Pp = prim[PPP];
  double vr = prim[URR];
  double vp = prim[UPP] * r;
  double vz = prim[UZZ];
  double v2 = vr * vr + vp * vp + vz * vz;
  double rhoe = Pp / (G - 1.);
  cons[DDD] = rho * dV;
  cons[SRR] = rho * vr * dV;
  cons[LLL] = r * rho * vp * dV;
  cons[SZZ] = rho * vz * dV;
  cons[TAU] = (.5 * rho * v2 + rhoe) * dV;
#pragma unroll
  for (int q = 5; q < NQ; ++q) cons[q] = prim[q] * cons[DDD];
}

// src/Cell/cell_conversion.c:57-103
template <int NQ>
DHD __forceinline__ void d_cons2prim(const double *cons, double *prim, double r, double dV,
                                            const DiskParams &P) {
  double G = P.gamma;
  double rho = cons[DDD] / dV;
  if (rho < P.rho_floor) rho = P.rho_floor;
  double Sr = cons[SRR] / dV;
  double Sp = cons[LLL] / dV / r;
  double Sz = cons[SZZ] / dV;
  double E = cons[TAU] / dV;
  double vr = Sr / rho;
  double vp = Sp / rho;
  double vz = Sz / rho;
  double KE = .5 * (Sr * vr + Sp * vp + Sz * vz);
  double v_magnitude = sqrt(2. * KE / rho);
  double rhoe = E - KE;
  double Pp = (G - 1.) * rhoe;
  if (Pp < P.cs_floor * P.cs_floor * rho / G) Pp = P.cs_floor * P.cs_floor * rho / G;
  if (Pp > P.cs_cap * P.cs_cap * rho / G) Pp = P.cs_cap * P.cs_cap * rho / G;
  if (v_magnitude > P.vel_cap) {
    vr = vr * P.vel_cap / v_magnitude;
    vp = vp * P.vel_cap / v_magnitude;
    vz = vz * P.vel_cap / v_magnitude;
  }
  prim[RHO] = rho;
  prim[PPP] = Pp;
  prim[URR] = vr;
  prim[UPP] = vp / r;
  prim[UZZ] = vz;
#pragma unroll
  for (int q = 5; q < NQ; ++q) prim[q] = cons[q] / cons[DDD];
}

// fgrav and gravMassForce of one mass at a cell, projected on r and phi,
// src/Cell/cell_source.c:11-13, 49-71.  epsn = pow(eps, n) is uniform over a
// launch (DiskScalars::epsn); xbh, ybh are rp*cos(pp), rp*sin(pp); cosp,
// sinp are cos and sin of the cell angle.
DHD __forceinline__ double d_fgrav(double M, double r, double epsn, int n) {
  if (n == 2) return (M * r / pow_p1p5(r * r + epsn));  // pow(r,1) = r, pow(r,2), pow(x,1.5)
  return (M * pow(r, n - 1.) / pow(pow(r, n) + epsn, 1. + 1. / n));
}
DHD __forceinline__ void d_grav_force(double M, double xbh, double ybh, double epsn, int order, double r,
                                      double cosp, double sinp, double *fr, double *fp) {
  double dx = r * cosp - xbh;
  double dy = r * sinp - ybh;
  double script_r = sqrt(dx * dx + dy * dy);
  double cosa = dx / script_r;
  double sina = dy / script_r;
  double cosap = cosa * cosp + sina * sinp;
  double sinap = sina * cosp - cosa * sinp;
  double f1 = -d_fgrav(M, script_r, epsn, order);
  *fr = cosap * f1;
  *fp = sinap * f1;
}

// The geometric, gravity and frame terms of cell_add_src
// (src/Cell/cell_source.c:117-173), added to u in the reference's order.
// pr: the cell's primitives, (cphi, sphi): cos and sin of its centre angle,
// zc its centre height, (w_a, rdrOm_a, dtOm_a) the frame rotation at r.
// Returns F_phi of the two masses in *Fp_out.
DHD __forceinline__ void add_geom_grav_src(double *u, const double *pr, double r, double dV, double dt, double cphi,
                                           double sphi, double zc, double w_a, double rdrOm_a, double dtOm_a,
                                           const DiskParams &P, const DiskScalars &sc, double *Fp_out) {
  double rho = pr[RHO], Pp = pr[PPP], vr = pr[URR], vz = pr[UZZ];
  double vp = pr[UPP] * r;
  double R = sqrt(r * r + zc * zc);
  double sint, cost, gravdist;
  if (P.grav2d == 1) {
    sint = 1.0;
    cost = 0.0;
    gravdist = r;
  } else {
    sint = r / R;
    cost = zc / R;
    gravdist = R;
  }
  double fr, fp, Fr = 0.0, Fp = 0.0;
  d_grav_force(sc.gm[0].M, sc.xbh[0], sc.ybh[0], sc.epsn, P.phi_order, gravdist, cphi, sphi, &fr, &fp);
  Fr += fr;
  Fp += fp;
  d_grav_force(sc.gm[1].M, sc.xbh[1], sc.ybh[1], sc.epsn, P.phi_order, gravdist, cphi, sphi, &fr, &fp);
  Fr += fr;
  Fp += fp;
  double F_centrifugal_r = w_a * w_a / r;
  double F_euler_phi = -(r * dtOm_a + vr * rdrOm_a);
  u[SRR] += dt * dV * (rho * vp * vp + Pp) / r;
  u[SRR] += dt * dV * rho * (Fr * sint + F_centrifugal_r);
  u[LLL] += dt * dV * rho * (Fp - r * dtOm_a) * r;
  u[SZZ] += dt * dV * rho * Fr * cost;
  u[TAU] += dt * dV * rho * ((Fr * sint + F_centrifugal_r) * vr + Fr * vz * cost + (Fp + F_euler_phi) * vp);
  *Fp_out = Fp;
}

// CFL time step of one cell, src/Cell/cell_mindt.c:44-90: pr its
// primitives, r its centre radius, dx its smallest extent, w the mesh
// velocity, tc its centre angle (read only by the alpha viscosity).
template <int FEAT = F_ALL>
DHD __forceinline__ double cell_dt_value(const DiskParams &P, const DiskScalars &sc, const double *pr, double r,
                                         double dx, double w, double tc) {
  double Pp = pr[PPP];
  double rho = pr[RHO];
  double vp = pr[UPP] * r - w;
  double vr = pr[URR];
  double vz = pr[UZZ];
  double cs = sqrt(P.gamma * Pp / rho);
  double cf2 = cs * cs;
  double a = sqrt(cf2) + sqrt(vr * vr + vp * vp + vz * vz);
  double d = P.cfl * dx / a;
  if ((FEAT & F_VISC) && P.visc > 0.0) {
    double nu;
    if (P.visc_const == 1) {
      nu = P.visc;
    } else if (P.initial_data_type == 0) {
      nu = P.visc * P.gamma * Pp / rho * pow(fabs(r * cos(tc)), 2.0);
      if (r * cos(tc) > 20.) nu = 0.000000001;
    } else {
      double ct, st;
      sincos(tc, &st, &ct);
      double d0 = d_grav_dist(sc.xbh[0], sc.ybh[0], r, ct, st, 0.);
      double d1 = d_grav_dist(sc.xbh[1], sc.ybh[1], r, ct, st, 0.);
      double eps = P.g_eps;
      nu = P.visc * Pp / rho / sqrt(pow_m1p5(d0 * d0 + eps * eps) * sc.gm[0].M + pow_m1p5(d1 * d1 + eps * eps) * sc.gm[1].M);
    }
    double dt_visc = .25 * dx * dx / nu;
    d = d / (1. + d / dt_visc);
  }
  return d;
}

// Flux of one face, src/Riemann/Riemann_routines.c:14-397.  The caller has
// reconstructed primL/primR (riemann_setup_*); n is the face normal
// (0: r, 1: phi, 2: z), w the face velocity.  On return X[q] = dt*dA*F[q]
// and, with viscosity, Y = dt*dA*{Fvisc[SRR], Fvisc[LLL], Fvisc[TAU]} --
// the quantities riemann_AddFlux adds to the two cells.
struct ViscIn {   // what riemann_visc_flux reads besides the reconstructed states
  double cm, r_cell_L, r_cell_R;
  const double *prim, *grad, *gradp;     // q-major fields (read only where needed)
  long long stride, cL, cR;
  double tL, tR, dL, dR;                 // tiph (wrapped), dphi of the two cells
  double rdr_om;                         // sim_rdrOm_a at the face radius
};


template <int NQ>
DHD __forceinline__ void d_set_flux(const double *prim, double *F, double r, int n0, int n1, int n2,
                                           double G) {  // Riemann_routines.c:157-191
  double rho = prim[RHO];
  double Pp = prim[PPP];
  double vr = prim[URR];
  double vp = prim[UPP] * r;
  double vz = prim[UZZ];
  double vn = vr * n0 + vp * n1 + vz * n2;
  double rhoe = Pp / (G - 1.);
  double v2 = vr * vr + vp * vp + vz * vz;
  F[DDD] = rho * vn;
  F[SRR] = rho * vr * vn + Pp * n0;
  F[LLL] = r * (rho * vp * vn + Pp * n1);
  F[SZZ] = rho * vz * vn + Pp * n2;
  F[TAU] = (.5 * rho * v2 + rhoe + Pp) * vn;
#pragma unroll
  for (int q = 5; q < NQ; ++q) F[q] = prim[q] * F[DDD];
}

// PART 0: the whole flux; 1: the Riemann increments X only; 2: the viscous
// increments Y only (the two parts share no intermediate but dt*dA).  FEAT:
// the solvers and the viscous flux this variant keeps (F_HLL | F_HLLC both:
// the disk's RiemannSolver picks at run time).
template <int NQ, int PART = 0, int FEAT = F_ALL>
DHD __forceinline__ int d_face_flux(const double *primL, const double *primR, int dir, double r,
                                           double dA, double w, double dt, const DiskParams &P,
                                           const DiskScalars &sc, const ViscIn *vi, double *X, double *Y) {
  const GravState *gm = sc.gm;
  const double G = P.gamma;
  const int n0 = dir == 0, n1 = dir == 1, n2 = dir == 2;
  double dtdA = dt * dA;
  if (PART != 2) {
    // riemann_set_vel, :30-55
    double vnL, cf2L, vnR, cf2R;
    {
      double Pp = primL[PPP], rho = primL[RHO], vr = primL[URR], vp = r * primL[UPP], vz = primL[UZZ];
      vnL = vr * n0 + vp * n1 + vz * n2;
      cf2L = G * (Pp / rho);
    }
    double Sl = vnL - sqrt(cf2L);
    double Sr = vnL + sqrt(cf2L);
    {
      double Pp = primR[PPP], rho = primR[RHO], vr = primR[URR], vp = r * primR[UPP], vz = primR[UZZ];
      vnR = vr * n0 + vp * n1 + vz * n2;
      cf2R = G * (Pp / rho);
    }
    if (Sl > vnR - sqrt(cf2R)) Sl = vnR - sqrt(cf2R);
    if (Sr < vnR + sqrt(cf2R)) Sr = vnR + sqrt(cf2R);
    double Ss = (primR[RHO] * vnR * (Sr - vnR) - primL[RHO] * vnL * (Sl - vnL) + primL[PPP] - primR[PPP]) /
                (primR[RHO] * (Sr - vnR) - primL[RHO] * (Sl - vnL));
    // riemann_set_state, :59-71
    int state;
    if (w < Sl)
      state = 0;  // LEFT
    else if (w > Sr)
      state = 3;  // RIGHT
    else if (w < Ss)
      state = 1;  // LEFTSTAR
    else
      state = 2;  // RIGHTSTAR
    double F[NQ];
    if (state == 0 || state == 3) {
      const double *pr = state == 0 ? primL : primR;
      double Fk[NQ], Uk[NQ];
      d_set_flux<NQ>(pr, Fk, r, n0, n1, n2, G);
      d_prim2cons<NQ>(pr, Uk, r, 1.0, G, P.NC);
#pragma unroll
      for (int q = 0; q < NQ; ++q) F[q] = Fk[q] - w * Uk[q];
    } else if ((FEAT & F_HLL) && (!(FEAT & F_HLLC) || P.riemann == HLL_SOLVER)) {  // Riemann_routines.c:74-85
      double FL[NQ], FR[NQ], UL[NQ], UR[NQ];
      d_set_flux<NQ>(primL, FL, r, n0, n1, n2, G);
      d_set_flux<NQ>(primR, FR, r, n0, n1, n2, G);
      d_prim2cons<NQ>(primL, UL, r, 1.0, G, P.NC);
      d_prim2cons<NQ>(primR, UR, r, 1.0, G, P.NC);
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        double Ustar = (Sr * UR[q] - Sl * UL[q] + FL[q] - FR[q]) / (Sr - Sl);
        double Fstar = (Sr * FL[q] - Sl * FR[q] - Sr * Sl * (UL[q] - UR[q])) / (Sr - Sl);
        F[q] = Fstar - w * Ustar;
      }
    } else {  // HLLC, Riemann_routines.c:87-154
      const double *prim = state == 1 ? primL : primR;
      double Sk = state == 1 ? Sl : Sr;
      double Fk[NQ], Uk[NQ], Ustar[NQ];
      d_set_flux<NQ>(prim, Fk, r, n0, n1, n2, G);
      d_prim2cons<NQ>(prim, Uk, r, 1.0, G, P.NC);
      double rho = prim[RHO];
      double vr = prim[URR];
      double vp = prim[UPP] * r;
      double vz = prim[UZZ];
      double Pp = prim[PPP];
      double v2 = vr * vr + vp * vp + vz * vz;
      double vn = vr * n0 + vp * n1 + vz * n2;
      double rhoe = Pp / (G - 1.);
      double D = rho;
      double mr = rho * vr;
      double mp = rho * vp;
      double mz = rho * vz;
      double E_hydro = .5 * rho * v2 + rhoe;
      double Ps = rho * (Sk - vn) * (Ss - vn) + Pp;
      double Dstar = (Sk - vn) * D / (Sk - Ss);
      double Msn = Dstar * Ss;
      double Msr = (Sk - vn) * mr / (Sk - Ss);
      double Msp = (Sk - vn) * mp / (Sk - Ss);
      double Msz = (Sk - vn) * mz / (Sk - Ss);
      double Estar = ((Sk - vn) * E_hydro + Ps * Ss - Pp * vn) / (Sk - Ss);
      double mn = Msr * n0 + Msp * n1 + Msz * n2;
      Msr += n0 * (Msn - mn);
      Msp += n1 * (Msn - mn);
      Msz += n2 * (Msn - mn);
      Ustar[DDD] = Dstar;
      Ustar[SRR] = Msr;
      Ustar[LLL] = r * Msp;
      Ustar[SZZ] = Msz;
      Ustar[TAU] = Estar;
#pragma unroll
      for (int q = 5; q < NQ; ++q) Ustar[q] = prim[q] * Ustar[DDD];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        double Fstar = Fk[q] + Sk * (Ustar[q] - Uk[q]);
        F[q] = Fstar - w * Ustar[q];
      }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) X[q] = dtdA * F[q];
  }
  if (PART != 1 && (FEAT & F_VISC) && P.visc > 0.0) {  // riemann_visc_flux, :193-311
    // only the entries the flux formulas use are formed: Avg for all five
    // primitives, the r and phi gradients of v_r and omega
    double Avg[5], Grr[2], Gph[2];
    double tiph = vi->cm;
    double M0 = gm[0].M, M1 = gm[1].M;
    double st, ct;
    sincos(tiph, &st, &ct);
    double dist_bh0 = d_grav_dist(sc.xbh[0], sc.ybh[0], r, ct, st, 0.);
    double dist_bh1 = d_grav_dist(sc.xbh[1], sc.ybh[1], r, ct, st, 0.);
    const double *pr = vi->prim, *gp = vi->gradp, *gr = vi->grad;
    const long long S = vi->stride, cL = vi->cL, cR = vi->cR;
    if (n0 == 1) {
      double pL = vi->tL - .5 * vi->dL;
      double pR = vi->tR - .5 * vi->dR;
      double dpL = vi->cm - pL;
      double dpR = -vi->cm + pR;
      dpL = wrap_pm(dpL);
      dpR = wrap_pm(dpR);
      double deltaL = r - vi->r_cell_L;
      double deltaR = vi->r_cell_R - r;
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        double WL = pr[q * S + cL] + dpL * gp[q * S + cL];
        double WR = pr[q * S + cR] - dpR * gp[q * S + cR];
        Avg[q] = 0.5 * (WL + WR);
        if (q == URR || q == UPP) {
          Grr[q - URR] = (WR - WL) / (deltaR + deltaL);
          Gph[q - URR] = .5 * (gp[q * S + cL] + gp[q * S + cR]) / r;
        }
      }
    } else if (n1 == 1) {
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        double vL = pr[q * S + cL], vR = pr[q * S + cR];
        Avg[q] = 0.5 * (vR + vL);
        if (q == URR || q == UPP) {
          Grr[q - URR] = .5 * (gr[q * S + cL] + gr[q * S + cR]);
          Gph[q - URR] = (vR - vL) / (vi->tR - vi->tL) / r;
        }
      }
    } else {
#pragma unroll
      for (int q = 0; q < 5; ++q) Avg[q] = .5 * (primL[q] + primR[q]);
      Grr[0] = Grr[1] = Gph[0] = Gph[1] = 0.0;  // z faces: the flux formulas are not reached
    }
    double nu;
    double alpha = P.visc;
    if (P.visc_const == 1) {
      nu = alpha;
    } else if (P.initial_data_type == 0) {  // SHEAR
      nu = alpha * G * Avg[PPP] / Avg[RHO] * pow(fabs(r * ct), 2.0);
      if (r * ct > 20.) nu = 0.0;
    } else {
      double eps = P.g_eps;
      nu = alpha * Avg[PPP] / Avg[RHO] /
           sqrt(pow_m1p5(dist_bh0 * dist_bh0 + eps * eps) * M0 + pow_m1p5(dist_bh1 * dist_bh1 + eps * eps) * M1);
    }
    double rho = Avg[RHO], vr = Avg[URR], om = Avg[UPP], vz = Avg[UZZ];
    if (fabs(vz) > 1.e-12) return DERR_VISC_VZ;
    double Gr_vr = Grr[0], Gr_om = Grr[1];
    double Gp_vr = Gph[0], Gp_om = Gph[1];
    double rdr_om_cell = vi->rdr_om;
    double v_srr = 0.0, v_lll = 0.0, v_tau = 0.0;
    if (n0 == 1) {
      v_srr = -nu * rho * (r * Gr_vr - vr - r * r * Gp_om);
      v_lll = -nu * rho * (r * (r * Gr_om + rdr_om_cell) + r * Gp_vr);
      v_tau = -nu * rho * (vr * (Gr_vr - vr / r - r * Gp_om) + r * om * (Gp_vr + (rdr_om_cell + r * Gr_om)));
    }
    if (n1 == 1) {
      v_srr = -nu * rho * (r * (r * Gr_om + rdr_om_cell) + r * Gp_vr);
      v_lll = -nu * rho * (-r * Gr_vr + vr + r * r * Gp_om);
      v_tau = -nu * rho * (-r * om * (Gr_vr - vr / r - r * Gp_om) + vr * (Gp_vr + (rdr_om_cell + r * Gr_om)));
    }
    Y[0] = dtdA * v_srr;
    Y[1] = dtdA * v_lll;
    Y[2] = dtdA * v_tau;
  }
  return 0;
}

