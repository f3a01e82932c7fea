// Hand-written reverse-mode derivatives (vector-Jacobian products) of the
// per-cell and per-face functions of physics.h.  Each vjp_* takes the
// forward inputs (it recomputes what it needs) and the adjoint of the
// outputs, and ADDS J^T * (output adjoint) to the input adjoints.  Branches
// (limiters, Riemann states, floors, min/max of wave speeds) take the branch
// the forward took: the derivative of the piece the forward evaluated.
//
// Scope: NUM_Q = 5, the non-rotating analytic frame (w_analytic_type 0, so
// the frame terms rOm_a, rdrOm_a, dtOm_a vanish), no sink, no cooling.
// Every function compiles for the host too, where the tests compare it with
// central differences of the forward forms.
#pragma once
#include "physics.h"

// cons = prim2cons(prim; r, dV), src/Cell/cell_conversion.c:11-33
DHD __forceinline__ void vjp_prim2cons(const double *prim, double r, double dV, double G, const double *cb,
                                       double *pb) {
  const double rho = prim[RHO], vr = prim[URR], vp = prim[UPP] * r, vz = prim[UZZ];
  const double v2 = vr * vr + vp * vp + vz * vz;
  pb[RHO] += dV * (cb[DDD] + vr * cb[SRR] + r * vp * cb[LLL] + vz * cb[SZZ] + .5 * v2 * cb[TAU]);
  pb[PPP] += dV * cb[TAU] / (G - 1.);
  pb[URR] += dV * rho * (cb[SRR] + vr * cb[TAU]);
  pb[UPP] += r * dV * rho * (r * cb[LLL] + vp * cb[TAU]);
  pb[UZZ] += dV * rho * (cb[SZZ] + vz * cb[TAU]);
}

// prim = cons2prim(cons; r, dV) with its floors and caps,
// src/Cell/cell_conversion.c:57-103
DHD __forceinline__ void vjp_cons2prim(const double *cons, double r, double dV, const DiskParams &P,
                                       const double *ob, double *cb) {
  const double G = P.gamma;
  const double rho0 = cons[DDD] / dV;
  const bool rho_floored = rho0 < P.rho_floor;
  const double rho = rho_floored ? P.rho_floor : rho0;
  const double Sr = cons[SRR] / dV, Sp = cons[LLL] / dV / r, Sz = cons[SZZ] / dV;
  const double vr = Sr / rho, vp = Sp / rho, vz = Sz / rho;
  const double KE = .5 * (Sr * vr + Sp * vp + Sz * vz);
  const double vmag = sqrt(2. * KE / rho);
  const double E = cons[TAU] / dV;
  double Pp = (G - 1.) * (E - KE);
  int pclamp = 0;
  if (Pp < P.cs_floor * P.cs_floor * rho / G) {
    Pp = P.cs_floor * P.cs_floor * rho / G;
    pclamp = 1;
  }
  if (Pp > P.cs_cap * P.cs_cap * rho / G) {
    Pp = P.cs_cap * P.cs_cap * rho / G;
    pclamp = 2;
  }
  const bool capped = vmag > P.vel_cap;
  // reverse
  double rhob = ob[RHO];
  const double vrf = ob[URR], vpf = ob[UPP] / r, vzf = ob[UZZ];
  double vrb = vrf, vpb = vpf, vzb = vzf, vmb = 0.0;
  if (capped) {
    const double s = P.vel_cap / vmag;
    vrb = s * vrf;
    vpb = s * vpf;
    vzb = s * vzf;
    vmb = -(vr * vrf + vp * vpf + vz * vzf) * P.vel_cap / (vmag * vmag);
  }
  double rhoeb = 0.0;
  if (pclamp == 1)
    rhob += ob[PPP] * P.cs_floor * P.cs_floor / G;
  else if (pclamp == 2)
    rhob += ob[PPP] * P.cs_cap * P.cs_cap / G;
  else
    rhoeb = ob[PPP] * (G - 1.);
  const double Eb = rhoeb;
  double KEb = -rhoeb;
  if (vmb != 0.0 && vmag > 0.0) {
    KEb += vmb / (rho * vmag);
    rhob += -vmb * KE / (rho * rho * vmag);
  }
  double Srb = .5 * vr * KEb, Spb = .5 * vp * KEb, Szb = .5 * vz * KEb;
  vrb += .5 * Sr * KEb;
  vpb += .5 * Sp * KEb;
  vzb += .5 * Sz * KEb;
  Srb += vrb / rho;
  Spb += vpb / rho;
  Szb += vzb / rho;
  rhob += -(vr * vrb + vp * vpb + vz * vzb) / rho;
  if (!rho_floored) cb[DDD] += rhob / dV;
  cb[SRR] += Srb / dV;
  cb[LLL] += Spb / dV / r;
  cb[SZZ] += Szb / dV;
  cb[TAU] += Eb / dV;
}

// F = physical flux of one state, Riemann_routines.c:157-191
DHD __forceinline__ void vjp_set_flux(const double *prim, double r, int n0, int n1, int n2, double G,
                                      const double *Fb, double *pb) {
  const double rho = prim[RHO], Pp = prim[PPP], vr = prim[URR], vp = prim[UPP] * r, vz = prim[UZZ];
  const double vn = vr * n0 + vp * n1 + vz * n2;
  const double rhoe = Pp / (G - 1.);
  const double v2 = vr * vr + vp * vp + vz * vz;
  const double vnb = rho * Fb[DDD] + rho * vr * Fb[SRR] + r * rho * vp * Fb[LLL] + rho * vz * Fb[SZZ] +
                     (.5 * rho * v2 + rhoe + Pp) * Fb[TAU];
  pb[RHO] += vn * Fb[DDD] + vr * vn * Fb[SRR] + r * vp * vn * Fb[LLL] + vz * vn * Fb[SZZ] + .5 * v2 * vn * Fb[TAU];
  pb[PPP] += n0 * Fb[SRR] + r * n1 * Fb[LLL] + n2 * Fb[SZZ] + (1. / (G - 1.) + 1.) * vn * Fb[TAU];
  pb[URR] += rho * vn * Fb[SRR] + rho * vr * vn * Fb[TAU] + n0 * vnb;
  pb[UPP] += r * (r * rho * vn * Fb[LLL] + rho * vp * vn * Fb[TAU] + n1 * vnb);
  pb[UZZ] += rho * vn * Fb[SZZ] + rho * vz * vn * Fb[TAU] + n2 * vnb;
}

// F = the HLL / HLLC flux through a face moving at w, before the dt*dA
// factor (Riemann_routines.c:14-154, the flux select of :332-371).  Returns
// the forward F in Fout.
DHD __forceinline__ void vjp_riemann(const double *pL, const double *pR, int dir, double r, double w,
                                     const DiskParams &P, const double *Fb, double *pLb, double *pRb, double *Fout) {
  const double G = P.gamma;
  const int n0 = dir == 0, n1 = dir == 1, n2 = dir == 2;
  const double vnL = pL[URR] * n0 + r * pL[UPP] * n1 + pL[UZZ] * n2;
  const double vnR = pR[URR] * n0 + r * pR[UPP] * n1 + pR[UZZ] * n2;
  const double cL = sqrt(G * (pL[PPP] / pL[RHO])), cR = sqrt(G * (pR[PPP] / pR[RHO]));
  double Sl = vnL - cL, Sr = vnL + cL;
  const bool slR = Sl > vnR - cR;
  if (slR) Sl = vnR - cR;
  const bool srR = Sr < vnR + cR;
  if (srR) Sr = vnR + cR;
  const double Nn = pR[RHO] * vnR * (Sr - vnR) - pL[RHO] * vnL * (Sl - vnL) + pL[PPP] - pR[PPP];
  const double Dn = pR[RHO] * (Sr - vnR) - pL[RHO] * (Sl - vnL);
  const double Ss = Nn / Dn;
  const int state = w < Sl ? 0 : (w > Sr ? 3 : (w < Ss ? 1 : 2));
  double Slb = 0.0, Srb = 0.0, Ssb = 0.0, vnLb = 0.0, vnRb = 0.0, cLb = 0.0, cRb = 0.0;
  if (state == 0 || state == 3) {
    const double *pr = state == 0 ? pL : pR;
    double *pb = state == 0 ? pLb : pRb;
    double Fk[5], Uk[5], Ub[5];
    d_set_flux<5>(pr, Fk, r, n0, n1, n2, G);
    d_prim2cons<5>(pr, Uk, r, 1.0, G, 5);
    for (int q = 0; q < 5; ++q) {
      Fout[q] = Fk[q] - w * Uk[q];
      Ub[q] = -w * Fb[q];
    }
    vjp_set_flux(pr, r, n0, n1, n2, G, Fb, pb);
    vjp_prim2cons(pr, r, 1.0, G, Ub, pb);
  } else if (P.riemann == HLL_SOLVER) {
    double FL[5], FR[5], UL[5], UR[5];
    d_set_flux<5>(pL, FL, r, n0, n1, n2, G);
    d_set_flux<5>(pR, FR, r, n0, n1, n2, G);
    d_prim2cons<5>(pL, UL, r, 1.0, G, 5);
    d_prim2cons<5>(pR, UR, r, 1.0, G, 5);
    double FLb[5], FRb[5], ULb[5], URb[5];
    const double d = Sr - Sl;
    for (int q = 0; q < 5; ++q) {
      const double Us = (Sr * UR[q] - Sl * UL[q] + FL[q] - FR[q]) / d;
      const double Fs = (Sr * FL[q] - Sl * FR[q] - Sr * Sl * (UL[q] - UR[q])) / d;
      Fout[q] = Fs - w * Us;
      const double Fsb = Fb[q], Usb = -w * Fb[q];
      URb[q] = Sr / d * Usb + Sr * Sl / d * Fsb;
      ULb[q] = -Sl / d * Usb - Sr * Sl / d * Fsb;
      FLb[q] = Usb / d + Sr / d * Fsb;
      FRb[q] = -Usb / d - Sl / d * Fsb;
      Srb += (UR[q] - Us) / d * Usb + (FL[q] - Sl * (UL[q] - UR[q]) - Fs) / d * Fsb;
      Slb += (Us - UL[q]) / d * Usb + (-FR[q] - Sr * (UL[q] - UR[q]) + Fs) / d * Fsb;
    }
    vjp_set_flux(pL, r, n0, n1, n2, G, FLb, pLb);
    vjp_set_flux(pR, r, n0, n1, n2, G, FRb, pRb);
    vjp_prim2cons(pL, r, 1.0, G, ULb, pLb);
    vjp_prim2cons(pR, r, 1.0, G, URb, pRb);
  } else {  // HLLC star state of the side the face velocity falls on
    const double *pr = state == 1 ? pL : pR;
    double *pb = state == 1 ? pLb : pRb;
    const double Sk = state == 1 ? Sl : Sr;
    double Fk[5], Uk[5];
    d_set_flux<5>(pr, Fk, r, n0, n1, n2, G);
    d_prim2cons<5>(pr, Uk, r, 1.0, G, 5);
    const double rho = pr[RHO], vr = pr[URR], vp = pr[UPP] * r, vz = pr[UZZ], Pp = pr[PPP];
    const double v2 = vr * vr + vp * vp + vz * vz;
    const double vn = vr * n0 + vp * n1 + vz * n2;
    const double rhoe = Pp / (G - 1.);
    const double mr = rho * vr, mp = rho * vp, mz = rho * vz;
    const double Eh = .5 * rho * v2 + rhoe;
    const double a = Sk - vn, b = Sk - Ss;
    const double Ps = rho * a * (Ss - vn) + Pp;
    const double Dstar = a * rho / b;
    const double Msn = Dstar * Ss;
    const double Msr = a * mr / b, Msp = a * mp / b, Msz = a * mz / b;
    const double Estar = (a * Eh + Ps * Ss - Pp * vn) / b;
    const double mn = Msr * n0 + Msp * n1 + Msz * n2;
    double Us[5];
    Us[DDD] = Dstar;
    Us[SRR] = Msr + n0 * (Msn - mn);
    Us[LLL] = r * (Msp + n1 * (Msn - mn));
    Us[SZZ] = Msz + n2 * (Msn - mn);
    Us[TAU] = Estar;
    double Usb[5], Ukb[5];
    double Skb = 0.0;
    for (int q = 0; q < 5; ++q) {
      Fout[q] = Fk[q] + Sk * (Us[q] - Uk[q]) - w * Us[q];
      Usb[q] = (Sk - w) * Fb[q];
      Ukb[q] = -Sk * Fb[q];
      Skb += (Us[q] - Uk[q]) * Fb[q];
    }
    vjp_set_flux(pr, r, n0, n1, n2, G, Fb, pb);
    vjp_prim2cons(pr, r, 1.0, G, Ukb, pb);
    double Dstarb = Usb[DDD], Estarb = Usb[TAU];
    const double Msr1b = Usb[SRR], Msp1b = r * Usb[LLL], Msz1b = Usb[SZZ];
    const double t = n0 * Msr1b + n1 * Msp1b + n2 * Msz1b;
    const double Msnb = t, mnb = -t;
    double Msrb = Msr1b + n0 * mnb, Mspb = Msp1b + n1 * mnb, Mszb = Msz1b + n2 * mnb;
    Dstarb += Msnb * Ss;
    Ssb += Msnb * Dstar;
    const double numb = Estarb / b;
    double bb = -Estarb * Estar / b;
    double ab = Eh * numb;
    const double Ehb = a * numb;
    const double Psb = Ss * numb;
    Ssb += Ps * numb;
    double Ppb = -vn * numb;
    double vnb = -Pp * numb;
    double rhob = a * (Ss - vn) * Psb;
    ab += rho * (Ss - vn) * Psb;
    Ssb += rho * a * Psb;
    vnb += -rho * a * Psb;
    Ppb += Psb;
    ab += mz / b * Mszb + mp / b * Mspb + mr / b * Msrb + rho / b * Dstarb;
    const double mzb = a / b * Mszb, mpb = a / b * Mspb, mrb = a / b * Msrb;
    rhob += a / b * Dstarb;
    bb += -(Msz * Mszb + Msp * Mspb + Msr * Msrb + Dstar * Dstarb) / b;
    Skb += ab + bb;
    vnb -= ab;
    Ssb -= bb;
    rhob += .5 * v2 * Ehb + vr * mrb + vp * mpb + vz * mzb;
    const double vrb = rho * vr * Ehb + rho * mrb + n0 * vnb;
    const double vpb = rho * vp * Ehb + rho * mpb + n1 * vnb;
    const double vzb = rho * vz * Ehb + rho * mzb + n2 * vnb;
    Ppb += Ehb / (G - 1.);
    pb[RHO] += rhob;
    pb[PPP] += Ppb;
    pb[URR] += vrb;
    pb[UPP] += r * vpb;
    pb[UZZ] += vzb;
    if (state == 1)
      Slb += Skb;
    else
      Srb += Skb;
  }
  if (Ssb != 0.0) {
    const double Nnb = Ssb / Dn, Dnb = -Ssb * Ss / Dn;
    pRb[RHO] += vnR * (Sr - vnR) * Nnb + (Sr - vnR) * Dnb;
    vnRb += pR[RHO] * (Sr - 2. * vnR) * Nnb - pR[RHO] * Dnb;
    Srb += pR[RHO] * vnR * Nnb + pR[RHO] * Dnb;
    pLb[RHO] += -vnL * (Sl - vnL) * Nnb - (Sl - vnL) * Dnb;
    vnLb += -pL[RHO] * (Sl - 2. * vnL) * Nnb + pL[RHO] * Dnb;
    Slb += -pL[RHO] * vnL * Nnb - pL[RHO] * Dnb;
    pLb[PPP] += Nnb;
    pRb[PPP] -= Nnb;
  }
  if (slR) {
    vnRb += Slb;
    cRb -= Slb;
  } else {
    vnLb += Slb;
    cLb -= Slb;
  }
  if (srR) {
    vnRb += Srb;
    cRb += Srb;
  } else {
    vnLb += Srb;
    cLb += Srb;
  }
  if (cL > 0.0) {
    const double c2b = cLb / (2. * cL);
    pLb[PPP] += G / pL[RHO] * c2b;
    pLb[RHO] -= G * pL[PPP] / (pL[RHO] * pL[RHO]) * c2b;
  }
  if (cR > 0.0) {
    const double c2b = cRb / (2. * cR);
    pRb[PPP] += G / pR[RHO] * c2b;
    pRb[RHO] -= G * pR[PPP] / (pR[RHO] * pR[RHO]) * c2b;
  }
  pLb[URR] += n0 * vnLb;
  pLb[UPP] += r * n1 * vnLb;
  pLb[UZZ] += n2 * vnLb;
  pRb[URR] += n0 * vnRb;
  pRb[UPP] += r * n1 * vnRb;
  pRb[UZZ] += n2 * vnRb;
}

// adjoints of the launch-uniform scalars a substep reads (dt, alpha, the
// masses and their positions) and the loss: what a launch reduces per disk
struct ScalarAdj {
  double dt, alpha, M[2], xbh[2], ybh[2], loss;
};

// nu of the alpha viscosity at (r, angle), given P/rho
// (Riemann_routines.c:259-263, cell_mindt.c:84), with what its VJP reuses:
// the angle's cos / sin, the point (x, y), each mass's offset, s = d^2 +
// eps^2 and s^-1.5, Q = sum M s^-1.5 and sqrt(Q).  The same operations as
// the forward's, so nu has its bits.
struct ViscAt {
  double ct, st, x, y, nu, Q, sq;
  double dx[2], dy[2], s[2], pm[2];
};
DHD __forceinline__ void visc_at_cs(const DiskParams &P, const DiskScalars &sc, double r, double ct, double st,
                                    double Pp, double rho, ViscAt &V) {
  V.ct = ct;
  V.st = st;
  if (P.visc_const == 1) {
    V.nu = P.visc;
    return;
  }
  const double eps = P.g_eps;
  V.x = r * ct;
  V.y = r * st;
  for (int p = 0; p < 2; ++p) {
    V.dx[p] = V.x - sc.xbh[p];
    V.dy[p] = V.y - sc.ybh[p];
    const double d = sqrt(V.dx[p] * V.dx[p] + V.dy[p] * V.dy[p]);
    V.s[p] = d * d + eps * eps;
    V.pm[p] = pow_m1p5(V.s[p]);
  }
  V.Q = V.pm[0] * sc.gm[0].M + V.pm[1] * sc.gm[1].M;
  V.sq = sqrt(V.Q);
  V.nu = P.visc * Pp / rho / V.sq;
}
DHD __forceinline__ void visc_at(const DiskParams &P, const DiskScalars &sc, double r, double ang, double Pp,
                                 double rho, ViscAt &V) {
  double st, ct;
  sincos(ang, &st, &ct);
  visc_at_cs(P, sc, r, ct, st, Pp, rho, V);
}
DHD __forceinline__ double visc_nu(const DiskParams &P, const DiskScalars &sc, double r, double ct, double st,
                                   double Pp, double rho) {
  ViscAt V;
  visc_at_cs(P, sc, r, ct, st, Pp, rho, V);
  return V.nu;
}
// VJP of nu at V; angb (optional): adjoint of the angle whose cos / sin are V.ct / V.st
DHD __forceinline__ void vjp_visc_nu(const DiskParams &P, const DiskScalars &sc, const ViscAt &V, double Pp,
                                     double rho, double nub, double *Ppb, double *rhob, ScalarAdj &sb,
                                     double *angb = nullptr) {
  if (P.visc_const == 1) {
    sb.alpha += nub;
    return;
  }
  sb.alpha += Pp / rho / V.sq * nub;
  *Ppb += P.visc / (rho * V.sq) * nub;
  *rhob += -P.visc * Pp / (rho * rho * V.sq) * nub;
  const double Qb = -V.nu / (2. * V.Q) * nub;
  for (int p = 0; p < 2; ++p) {
    const double dx = V.dx[p], dy = V.dy[p], s = V.s[p];
    sb.M[p] += V.pm[p] * Qb;
    const double sbar = sc.gm[p].M * (-1.5) * V.pm[p] / s * Qb;
    // s = d^2 + eps^2 with d^2 = dx^2 + dy^2, (x, y) = r (cos, sin)(angle)
    sb.xbh[p] += -2. * dx * sbar;
    sb.ybh[p] += -2. * dy * sbar;
    if (angb) *angb += 2. * (dx * (-V.y) + dy * V.x) * sbar;
  }
}

// The three viscous flux components (Riemann_routines.c:286-301) from the
// face averages Avg[5], the r and phi gradients of (v_r, omega) Grr[2],
// Gph[2], for an r (n0) or phi (n1) face; V: nu at the face centre
// (visc_at with Avg's P and rho).
DHD __forceinline__ void visc_flux_v(const ViscAt &V, int n0, double r, double rdr, const double *Avg,
                                     const double *Grr, const double *Gph, double *v) {
  const double nu = V.nu;
  const double rho = Avg[RHO], vr = Avg[URR], om = Avg[UPP];
  const double K = -nu * rho;
  const double A1 = r * Grr[0] - vr - r * r * Gph[1];
  const double A2 = r * (r * Grr[1] + rdr) + r * Gph[0];
  const double B1 = Grr[0] - vr / r - r * Gph[1];
  const double B2 = Gph[0] + (rdr + r * Grr[1]);
  if (n0) {
    v[0] = K * A1;
    v[1] = K * A2;
    v[2] = K * (vr * B1 + r * om * B2);
  } else {
    v[0] = K * A2;
    v[1] = K * (-r * Grr[0] + vr + r * r * Gph[1]);
    v[2] = K * (-r * om * B1 + vr * B2);
  }
}
DHD __forceinline__ void vjp_visc_flux(const DiskParams &P, const DiskScalars &sc, const ViscAt &V, int n0, double r,
                                       double rdr, const double *Avg, const double *Grr, const double *Gph,
                                       const double *vb, double *Avgb, double *Grrb, double *Gphb, ScalarAdj &sb,
                                       double *cmb) {
  const double nu = V.nu;
  const double rho = Avg[RHO], vr = Avg[URR], om = Avg[UPP];
  const double K = -nu * rho;
  const double A1 = r * Grr[0] - vr - r * r * Gph[1];
  const double A2 = r * (r * Grr[1] + rdr) + r * Gph[0];
  const double B1 = Grr[0] - vr / r - r * Gph[1];
  const double B2 = Gph[0] + (rdr + r * Grr[1]);
  double Kb, A1b, A2b, B1b, B2b, vrb = 0.0, omb = 0.0;
  if (n0) {
    const double A3 = vr * B1 + r * om * B2;
    Kb = vb[0] * A1 + vb[1] * A2 + vb[2] * A3;
    A1b = K * vb[0];
    A2b = K * vb[1];
    const double A3b = K * vb[2];
    vrb += B1 * A3b;
    B1b = vr * A3b;
    omb += r * B2 * A3b;
    B2b = r * om * A3b;
  } else {
    const double A3 = -r * om * B1 + vr * B2;
    Kb = vb[0] * A2 + vb[1] * (-A1) + vb[2] * A3;
    A2b = K * vb[0];
    A1b = -K * vb[1];
    const double A3b = K * vb[2];
    omb += -r * B1 * A3b;
    B1b = -r * om * A3b;
    vrb += B2 * A3b;
    B2b = vr * A3b;
  }
  Grrb[0] += r * A1b + B1b;
  vrb += -A1b - B1b / r;
  Gphb[1] += -r * r * A1b - r * B1b;
  Grrb[1] += r * r * A2b + r * B2b;
  Gphb[0] += r * A2b + B2b;
  Avgb[URR] += vrb;
  Avgb[UPP] += omb;
  Avgb[RHO] += -nu * Kb;
  const double nub = -rho * Kb;
  vjp_visc_nu(P, sc, V, Avg[PPP], Avg[RHO], nub, &Avgb[PPP], &Avgb[RHO], sb, cmb);
}

// Gravity of one mass at a cell, projected on r and phi (cell_source.c:49-71).
DHD __forceinline__ void vjp_grav_force(double M, double xbh, double ybh, double epsn, int order, double r,
                                        double cosp, double sinp, double frb, double fpb, double *Mb, double *xbhb,
                                        double *ybhb, double *cospb, double *sinpb) {
  const double dx = r * cosp - xbh, dy = r * sinp - ybh;
  const double s = sqrt(dx * dx + dy * dy);
  const double cosa = dx / s, sina = dy / s;
  const double cosap = cosa * cosp + sina * sinp;
  const double sinap = sina * cosp - cosa * sinp;
  const double f1 = -d_fgrav(M, s, epsn, order);
  // f1 = -M g(s), g(s) = s^(n-1) / (s^n + eps^n)^(1 + 1/n)
  double g, gs;
  if (order == 2) {
    const double u = s * s + epsn;
    g = s / pow_p1p5(u);
    gs = (epsn - 2. * s * s) / (pow_p1p5(u) * u);
  } else {
    const double n = order;
    const double u = pow(s, n) + epsn;
    g = pow(s, n - 1.) / pow(u, 1. + 1. / n);
    gs = (n - 1.) * pow(s, n - 2.) / pow(u, 1. + 1. / n) - (n + 1.) * pow(s, 2. * n - 2.) / pow(u, 2. + 1. / n);
  }
  const double f1b = cosap * frb + sinap * fpb;
  const double cosapb = f1 * frb, sinapb = f1 * fpb;
  const double cosab = cosp * cosapb - sinp * sinapb;
  const double sinab = sinp * cosapb + cosp * sinapb;
  *Mb += -g * f1b;
  double sb = -M * gs * f1b;
  double dxb = cosab / s, dyb = sinab / s;
  sb += -(dx * cosab + dy * sinab) / (s * s);
  dxb += dx / s * sb;
  dyb += dy / s * sb;
  *xbhb += -dxb;
  *ybhb += -dyb;
  // cosp, sinp enter dx = r cosp - xbh, dy = r sinp - ybh and the projection
  *cospb += r * dxb + cosa * cosapb + sina * sinapb;
  *sinpb += r * dyb + sina * cosapb - cosa * sinapb;
}

// X = dt*dA*F and Y = dt*dA*(viscous SRR, LLL, TAU) of one face
// (d_face_flux) given the reconstructed states and, for viscosity, the face
// averages Avg, gradients Grr / Gph, face centre angle cm and rdrOm_a.
DHD __forceinline__ void vjp_face_flux(const DiskParams &P, const DiskScalars &sc, const double *pL, const double *pR,
                                       int dir, double r, double dA, double w, double dt, double cm, double rdr,
                                       const double *Avg, const double *Grr, const double *Gph, const double *Xb,
                                       const double *Yb, double *pLb, double *pRb, double *Avgb, double *Grrb,
                                       double *Gphb, ScalarAdj &sb, double *dAb = nullptr, double *cmb = nullptr) {
  const double dtdA = dt * dA;
  double Fb[5], F[5];
  for (int q = 0; q < 5; ++q) Fb[q] = dtdA * Xb[q];
  vjp_riemann(pL, pR, dir, r, w, P, Fb, pLb, pRb, F);
  double sdt = 0.0;
  for (int q = 0; q < 5; ++q) sdt += F[q] * Xb[q];
  if (P.visc > 0.0 && dir != 2) {
    double v[3], vb[3];
    ViscAt V;
    visc_at(P, sc, r, cm, Avg[PPP], Avg[RHO], V);
    visc_flux_v(V, dir == 0, r, rdr, Avg, Grr, Gph, v);
    for (int k = 0; k < 3; ++k) {
      vb[k] = dtdA * Yb[k];
      sdt += v[k] * Yb[k];
    }
    vjp_visc_flux(P, sc, V, dir == 0, r, rdr, Avg, Grr, Gph, vb, Avgb, Grrb, Gphb, sb, cmb);
  }
  sb.dt += dA * sdt;
  if (dAb) *dAb += dt * sdt;
}

// VJP of add_geom_grav_src (physics.h) with the frame terms of
// w_analytic_type 0 (zero), plus the loss term wl * dt*dV*rho*F_phi*r of
// an interior cell (wl = 0 elsewhere).  phi: cell-centre angle.
DHD __forceinline__ void vjp_cell_sources(const DiskParams &P, const DiskScalars &sc, const double *pr, double r,
                                          double dV, double dt, double phi, double zc, const double *ub, double wl,
                                          double *prb, ScalarAdj &sb, double *phib) {
  const double epsn = sc.epsn;
  double cphi, sphi;
  sincos(phi, &sphi, &cphi);
  const double rho = pr[RHO], Pp = pr[PPP], vr = pr[URR], vz = pr[UZZ], vp = pr[UPP] * r;
  const double R = sqrt(r * r + zc * zc);
  double sint = 1.0, cost = 0.0, gdist = r;
  if (P.grav2d != 1) {
    sint = r / R;
    cost = zc / R;
    gdist = R;
  }
  double Fr = 0.0, Fp = 0.0;
  for (int p = 0; p < 2; ++p) {
    double fr, fp;
    d_grav_force(sc.gm[p].M, sc.xbh[p], sc.ybh[p], epsn, P.phi_order, gdist, cphi, sphi, &fr, &fp);
    Fr += fr;
    Fp += fp;
  }
  const double uS = ub[SRR], uL = ub[LLL] + wl, uZ = ub[SZZ], uT = ub[TAU];
  const double TS = (rho * vp * vp + Pp) / r + rho * (Fr * sint);
  const double TL = rho * Fp * r, TZ = rho * Fr * cost;
  const double TT = rho * (Fr * sint * vr + Fr * vz * cost + Fp * vp);
  sb.dt += dV * (TS * uS + TL * uL + TZ * uZ + TT * uT);
  sb.loss += wl * dt * dV * TL;
  const double k = dt * dV;
  prb[RHO] += k * (vp * vp / r * uS + Fr * sint * uS + Fp * r * uL + Fr * cost * uZ +
                   (Fr * sint * vr + Fr * vz * cost + Fp * vp) * uT);
  prb[PPP] += k * uS / r;
  prb[UPP] += r * k * (2. * rho * vp / r * uS + rho * Fp * uT);
  prb[URR] += k * rho * Fr * sint * uT;
  prb[UZZ] += k * rho * Fr * cost * uT;
  const double Frb = k * rho * (sint * uS + cost * uZ + (sint * vr + vz * cost) * uT);
  const double Fpb = k * rho * (r * uL + vp * uT);
  double cb = 0.0, sbb = 0.0;
  for (int p = 0; p < 2; ++p)
    vjp_grav_force(sc.gm[p].M, sc.xbh[p], sc.ybh[p], epsn, P.phi_order, gdist, cphi, sphi, Frb, Fpb, &sb.M[p],
                   &sb.xbh[p], &sb.ybh[p], &cb, &sbb);
  *phib += -sphi * cb + cphi * sbb;
}

// VJP of cell_dt_value (physics.h) for constant or alpha-law viscosity.
DHD __forceinline__ void vjp_cell_dt(const DiskParams &P, const DiskScalars &sc, const double *pr, double r, double dx,
                                     double w, double tc, double db, double *prb, ScalarAdj &sb, double *tcb) {
  const double Pp = pr[PPP], rho = pr[RHO], vp = pr[UPP] * r - w, vr = pr[URR], vz = pr[UZZ];
  const double c = sqrt(P.gamma * Pp / rho);
  const double vm = sqrt(vr * vr + vp * vp + vz * vz);
  const double a = c + vm;
  const double d0 = P.cfl * dx / a;
  double d0b = db;
  double Ppb = 0.0, rhob = 0.0;
  if (P.visc > 0.0) {
    ViscAt V;
    visc_at(P, sc, r, tc, Pp, rho, V);
    const double nu = V.nu;
    const double tv = .25 * dx * dx / nu;
    const double den = 1. + d0 / tv;
    d0b = db / (den * den);
    const double tvb = db * d0 * d0 / (tv * tv * den * den);
    const double nub = -tvb * tv / nu;
    vjp_visc_nu(P, sc, V, Pp, rho, nub, &Ppb, &rhob, sb, tcb);
  }
  const double ab = -d0 / a * d0b;
  if (c > 0.0) {
    const double c2b = ab / (2. * c);
    Ppb += P.gamma / rho * c2b;
    rhob -= P.gamma * Pp / (rho * rho) * c2b;
  }
  prb[PPP] += Ppb;
  prb[RHO] += rhob;
  if (vm > 0.0) {
    prb[URR] += vr / vm * ab;
    prb[UPP] += r * vp / vm * ab;
    prb[UZZ] += vz / vm * ab;
  }
}

// phi slope of one quantity (cell_plm.c:141-153): which branch the limiter
// took and the weights of (v_L, v_c, v_R) in the result.
DHD __forceinline__ void plm_p_weights(double vL, double vc, double vR, double dpL, double dpR, double PLM,
                                       double *wL, double *wc, double *wR) {
  const double sL = (vc - vL) / dpL, sR = (vR - vc) / dpR, sM = (vR - vL) / (dpL + dpR);
  double s = PLM * sL;
  int br = 0;
  if (fabs(s) > PLM * fabs(sR)) {
    s = PLM * sR;
    br = 1;
  }
  if (fabs(s) > fabs(sM)) {
    s = sM;
    br = 2;
  }
  if (sL * sR < 0.0) br = 3;
  *wL = *wc = *wR = 0.0;
  if (br == 0) {
    *wc = PLM / dpL;
    *wL = -PLM / dpL;
  } else if (br == 1) {
    *wR = PLM / dpR;
    *wc = -PLM / dpR;
  } else if (br == 2) {
    *wR = 1. / (dpL + dpR);
    *wL = -1. / (dpL + dpR);
  }
}

// The two cells of a face: primitives, r (or z) slopes, phi slopes.
struct FacePair {
  double pL[5], pR[5], gL[5], gR[5], gpL[5], gpR[5];
};
// An r face (k_flux_rz): radius, the two half-widths, area, centre angle,
// the cells' wrapped edges and widths, rdrOm_a at the face.
struct RFaceGeo {
  double r, deltaL, deltaR, dA, cm, tL, dL, tR, dR, rdr;
};
// A phi face (phi_face_flux): centre radius, area, face velocity, the
// cells' wrapped edges and widths, rdrOm_a at the centre.
struct PhiFaceGeo {
  double r, dA, w, tL, tR, dL, dR, rdr;
};

DHD __forceinline__ void face_pair_arrays(const FacePair &c, double *pr, double *gr, double *gp) {
  for (int q = 0; q < 5; ++q) {
    pr[2 * q] = c.pL[q];
    pr[2 * q + 1] = c.pR[q];
    gr[2 * q] = c.gL[q];
    gr[2 * q + 1] = c.gR[q];
    gp[2 * q] = c.gpL[q];
    gp[2 * q + 1] = c.gpR[q];
  }
}

// r face: riemann_setup_rz (Riemann_setup.c:11-44) + the flux, exactly as
// k_flux_rz evaluates it, through d_face_flux.
DHD __forceinline__ void rface_flux(const DiskParams &P, const DiskScalars &sc, const FacePair &c, const RFaceGeo &g,
                                    double dt, double *X, double *Y) {
  const double dpL = wrap_pm(g.cm - (g.tL - .5 * g.dL));
  const double dpR = wrap_pm(-g.cm + (g.tR - .5 * g.dR));
  double primL[5], primR[5], pr[10], gr[10], gp[10];
  for (int q = 0; q < 5; ++q) {
    primL[q] = c.pL[q] + c.gL[q] * g.deltaL + c.gpL[q] * dpL;
    primR[q] = c.pR[q] - c.gR[q] * g.deltaR - c.gpR[q] * dpR;
  }
  face_pair_arrays(c, pr, gr, gp);
  ViscIn vi;
  vi.prim = pr;
  vi.grad = gr;
  vi.gradp = gp;
  vi.stride = 2;
  vi.cL = 0;
  vi.cR = 1;
  vi.tL = g.tL;
  vi.tR = g.tR;
  vi.dL = g.dL;
  vi.dR = g.dR;
  vi.cm = g.cm;
  vi.r_cell_L = g.r - g.deltaL;
  vi.r_cell_R = g.r + g.deltaR;
  vi.rdr_om = g.rdr;
  Y[0] = Y[1] = Y[2] = 0.0;
  d_face_flux<5>(primL, primR, 0, g.r, g.dA, 0.0, dt, P, sc, &vi, X, Y);
}
// The r face's VJP in two parts (the kernels run them as two launches to keep
// the register count down): the Riemann flux of the reconstructed states and
// the viscous flux of the phi-shifted states.  Each returns the adjoints of
// the quantities it read from the two cells, of the face area dA, of the
// offsets dpL, dpR (which move with the face centre and the cells' edges)
// and, for the viscous part, of the face centre cm.
struct RFaceHydroAdj {
  double pLb[5], pRb[5];  // of primL = p_L + g_L dL + gp_L dpL and primR = p_R - g_R dR - gp_R dpR
  double dAb, dpLb, dpRb;
};
DHD __forceinline__ void vjp_rface_hydro(const DiskParams &P, const FacePair &c, const RFaceGeo &g, double dt,
                                         const double *Xb, ScalarAdj &sb, RFaceHydroAdj &h) {
  const double dpL = wrap_pm(g.cm - (g.tL - .5 * g.dL));
  const double dpR = wrap_pm(-g.cm + (g.tR - .5 * g.dR));
  double primL[5], primR[5], Fb[5], F[5];
  for (int q = 0; q < 5; ++q) {
    primL[q] = c.pL[q] + c.gL[q] * g.deltaL + c.gpL[q] * dpL;
    primR[q] = c.pR[q] - c.gR[q] * g.deltaR - c.gpR[q] * dpR;
    Fb[q] = dt * g.dA * Xb[q];
    h.pLb[q] = h.pRb[q] = 0.0;
  }
  vjp_riemann(primL, primR, 0, g.r, 0.0, P, Fb, h.pLb, h.pRb, F);
  double sdt = 0.0;
  for (int q = 0; q < 5; ++q) sdt += F[q] * Xb[q];
  sb.dt += g.dA * sdt;
  h.dAb = dt * sdt;
  h.dpLb = h.dpRb = 0.0;
  for (int q = 0; q < 5; ++q) {
    h.dpLb += h.pLb[q] * c.gpL[q];
    h.dpRb -= h.pRb[q] * c.gpR[q];
  }
}
struct RFaceViscAdj {
  double WLb[5], WRb[5];  // of W_L = p_L + dpL gp_L and W_R = p_R - dpR gp_R
  double Gphb[2];         // of Gph = (gp_L + gp_R)/(2r) for v_r, omega
  double dAb, cmb, dpLb, dpRb;
};
DHD __forceinline__ void vjp_rface_visc(const DiskParams &P, const DiskScalars &sc, const FacePair &c,
                                        const RFaceGeo &g, double dt, const double *Yb, ScalarAdj &sb,
                                        RFaceViscAdj &v) {
  const double dpL = wrap_pm(g.cm - (g.tL - .5 * g.dL));
  const double dpR = wrap_pm(-g.cm + (g.tR - .5 * g.dR));
  double WL[5], WR[5], Avg[5], Grr[2], Gph[2];
  for (int q = 0; q < 5; ++q) {
    WL[q] = c.pL[q] + dpL * c.gpL[q];
    WR[q] = c.pR[q] - dpR * c.gpR[q];
    Avg[q] = 0.5 * (WL[q] + WR[q]);
  }
  const double dsum = ((g.r + g.deltaR) - g.r) + (g.r - (g.r - g.deltaL));
  for (int k = 0; k < 2; ++k) {
    Grr[k] = (WR[URR + k] - WL[URR + k]) / dsum;
    Gph[k] = .5 * (c.gpL[URR + k] + c.gpR[URR + k]) / g.r;
  }
  double vf[3], vb[3], Avgb[5] = {0, 0, 0, 0, 0}, Grrb[2] = {0, 0};
  ViscAt V;
  visc_at(P, sc, g.r, g.cm, Avg[PPP], Avg[RHO], V);
  visc_flux_v(V, 1, g.r, g.rdr, Avg, Grr, Gph, vf);
  double sdt = 0.0;
  for (int k = 0; k < 3; ++k) {
    vb[k] = dt * g.dA * Yb[k];
    sdt += vf[k] * Yb[k];
  }
  sb.dt += g.dA * sdt;
  v.dAb = dt * sdt;
  v.cmb = 0.0;
  v.Gphb[0] = v.Gphb[1] = 0.0;
  vjp_visc_flux(P, sc, V, 1, g.r, g.rdr, Avg, Grr, Gph, vb, Avgb, Grrb, v.Gphb, sb, &v.cmb);
  for (int q = 0; q < 5; ++q) v.WLb[q] = v.WRb[q] = .5 * Avgb[q];
  for (int k = 0; k < 2; ++k) {
    v.WLb[URR + k] -= Grrb[k] / dsum;
    v.WRb[URR + k] += Grrb[k] / dsum;
  }
  v.dpLb = v.dpRb = 0.0;
  for (int q = 0; q < 5; ++q) {
    v.dpLb += v.WLb[q] * c.gpL[q];
    v.dpRb -= v.WRb[q] * c.gpR[q];
  }
}

// gb accumulates the adjoints of the face area dA, centre cm and the cells'
// edges tL, tR (the mesh geometry, which moves with dt); both parts together
DHD __forceinline__ void vjp_rface_flux(const DiskParams &P, const DiskScalars &sc, const FacePair &c,
                                        const RFaceGeo &g, double dt, const double *Xb, const double *Yb,
                                        FacePair &cb, ScalarAdj &sb, RFaceGeo &gb) {
  const double dpL = wrap_pm(g.cm - (g.tL - .5 * g.dL));
  const double dpR = wrap_pm(-g.cm + (g.tR - .5 * g.dR));
  RFaceHydroAdj h;
  vjp_rface_hydro(P, c, g, dt, Xb, sb, h);
  double dpLb = h.dpLb, dpRb = h.dpRb;
  gb.dA += h.dAb;
  for (int q = 0; q < 5; ++q) {
    cb.pL[q] += h.pLb[q];
    cb.gL[q] += g.deltaL * h.pLb[q];
    cb.gpL[q] += dpL * h.pLb[q];
    cb.pR[q] += h.pRb[q];
    cb.gR[q] -= g.deltaR * h.pRb[q];
    cb.gpR[q] -= dpR * h.pRb[q];
  }
  if (P.visc > 0.0) {
    RFaceViscAdj v;
    vjp_rface_visc(P, sc, c, g, dt, Yb, sb, v);
    gb.dA += v.dAb;
    gb.cm += v.cmb;
    dpLb += v.dpLb;
    dpRb += v.dpRb;
    for (int q = 0; q < 5; ++q) {
      cb.pL[q] += v.WLb[q];
      cb.gpL[q] += dpL * v.WLb[q];
      cb.pR[q] += v.WRb[q];
      cb.gpR[q] -= dpR * v.WRb[q];
    }
    for (int k = 0; k < 2; ++k) {
      cb.gpL[URR + k] += .5 * v.Gphb[k] / g.r;
      cb.gpR[URR + k] += .5 * v.Gphb[k] / g.r;
    }
  }
  // dpL = cm - (tL - dL/2), dpR = (tR - dR/2) - cm
  gb.cm += dpLb - dpRb;
  gb.tL -= dpLb;
  gb.tR += dpRb;
}

// phi face (j, j+1): riemann_setup_p (Riemann_setup.c:46-80) + the flux, as
// phi_face_flux evaluates it.
DHD __forceinline__ void phiface_flux(const DiskParams &P, const DiskScalars &sc, const FacePair &c,
                                      const PhiFaceGeo &g, double dt, double *X, double *Y) {
  double primL[5], primR[5], pr[10], gr[10], gp[10];
  for (int q = 0; q < 5; ++q) {
    primL[q] = c.pL[q] + 0.5 * c.gpL[q] * g.dL;
    primR[q] = c.pR[q] - 0.5 * c.gpR[q] * g.dR;
  }
  face_pair_arrays(c, pr, gr, gp);
  ViscIn vi;
  vi.prim = pr;
  vi.grad = gr;
  vi.gradp = gp;
  vi.stride = 2;
  vi.cL = 0;
  vi.cR = 1;
  vi.tL = g.tL;
  vi.tR = g.tR;
  vi.dL = g.dL;
  vi.dR = g.dR;
  vi.cm = g.tL;
  vi.r_cell_L = g.r;
  vi.r_cell_R = g.r;
  vi.rdr_om = g.rdr;
  Y[0] = Y[1] = Y[2] = 0.0;
  d_face_flux<5>(primL, primR, 1, g.r, g.dA, g.w, dt, P, sc, &vi, X, Y);
}
// The phi face's VJP in two parts, as for the r face: the Riemann flux of
// primL = p_L + gp_L dL/2, primR = p_R - gp_R dR/2, and the viscous flux of
// the cell averages with the r slopes and the (unwrapped) phi difference;
// the viscosity is evaluated at the face's angle tL.
DHD __forceinline__ void vjp_phiface_hydro(const DiskParams &P, const FacePair &c, const PhiFaceGeo &g, double dt,
                                           const double *Xb, ScalarAdj &sb, double *pLb, double *pRb) {
  double primL[5], primR[5], Fb[5], F[5];
  for (int q = 0; q < 5; ++q) {
    primL[q] = c.pL[q] + 0.5 * c.gpL[q] * g.dL;
    primR[q] = c.pR[q] - 0.5 * c.gpR[q] * g.dR;
    Fb[q] = dt * g.dA * Xb[q];
    pLb[q] = pRb[q] = 0.0;
  }
  vjp_riemann(primL, primR, 1, g.r, g.w, P, Fb, pLb, pRb, F);
  double sdt = 0.0;
  for (int q = 0; q < 5; ++q) sdt += F[q] * Xb[q];
  sb.dt += g.dA * sdt;
}
struct PhiFaceViscAdj {
  double Avgb[5], Grrb[2], Gphb[2], tLb;
};
DHD __forceinline__ void vjp_phiface_visc(const DiskParams &P, const DiskScalars &sc, const FacePair &c,
                                          const PhiFaceGeo &g, double dt, const double *Yb, ScalarAdj &sb,
                                          PhiFaceViscAdj &v) {
  double Avg[5], Grr[2], Gph[2], vf[3], vb[3];
  for (int q = 0; q < 5; ++q) {
    Avg[q] = 0.5 * (c.pR[q] + c.pL[q]);
    v.Avgb[q] = 0.0;
  }
  for (int k = 0; k < 2; ++k) {
    Grr[k] = .5 * (c.gL[URR + k] + c.gR[URR + k]);
    Gph[k] = (c.pR[URR + k] - c.pL[URR + k]) / (g.tR - g.tL) / g.r;
    v.Grrb[k] = v.Gphb[k] = 0.0;
  }
  ViscAt V;
  visc_at(P, sc, g.r, g.tL, Avg[PPP], Avg[RHO], V);
  visc_flux_v(V, 0, g.r, g.rdr, Avg, Grr, Gph, vf);
  double sdt = 0.0;
  for (int k = 0; k < 3; ++k) {
    vb[k] = dt * g.dA * Yb[k];
    sdt += vf[k] * Yb[k];
  }
  sb.dt += g.dA * sdt;
  v.tLb = 0.0;
  vjp_visc_flux(P, sc, V, 0, g.r, g.rdr, Avg, Grr, Gph, vb, v.Avgb, v.Grrb, v.Gphb, sb, &v.tLb);
}

// *tLb: adjoint of the face's angle (the viscosity is evaluated at cm = tL);
// both parts together
DHD __forceinline__ void vjp_phiface_flux(const DiskParams &P, const DiskScalars &sc, const FacePair &c,
                                          const PhiFaceGeo &g, double dt, const double *Xb, const double *Yb,
                                          FacePair &cb, ScalarAdj &sb, double *tLb) {
  double pLb[5], pRb[5];
  vjp_phiface_hydro(P, c, g, dt, Xb, sb, pLb, pRb);
  for (int q = 0; q < 5; ++q) {
    cb.pL[q] += pLb[q];
    cb.gpL[q] += .5 * g.dL * pLb[q];
    cb.pR[q] += pRb[q];
    cb.gpR[q] -= .5 * g.dR * pRb[q];
  }
  if (P.visc > 0.0) {
    PhiFaceViscAdj v;
    vjp_phiface_visc(P, sc, c, g, dt, Yb, sb, v);
    const double dt_phi = (g.tR - g.tL) * g.r;
    for (int q = 0; q < 5; ++q) {
      cb.pL[q] += .5 * v.Avgb[q];
      cb.pR[q] += .5 * v.Avgb[q];
    }
    for (int k = 0; k < 2; ++k) {
      cb.gL[URR + k] += .5 * v.Grrb[k];
      cb.gR[URR + k] += .5 * v.Grrb[k];
      cb.pR[URR + k] += v.Gphb[k] / dt_phi;
      cb.pL[URR + k] -= v.Gphb[k] / dt_phi;
    }
    *tLb += v.tLb;
  }
}
