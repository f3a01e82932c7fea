// Per-cell operations of the substep shared by the fused update (k_update,
// kernels.hip) and the unit-level class-API kernels (unit.hip): the source
// terms of cell_add_src, cell_add_visc_src and cell_add_split_fictitious
// (src/Cell/cell_source.c:73-277, src/Cell/cell_add_split_fictitious.c:11-44),
// each in the reference's expression order.
#pragma once
#include "device_common.h"

// cell_add_src for one cell (src/Cell/cell_source.c:117-203): the geometric,
// gravity and frame terms, then the accretion sink and the cooling.  cphi,
// sphi: cos and sin of the cell's centre angle; zc its centre height; rot
// the frame rotation of its annulus.  The sink's accretion of interior cells
// is added to *md0 / *md1; *cool gets the cell's cooling rate (when cooling).
// Returns F_phi of the two masses (the torque term of the LLL source).
template <int FEAT>
__device__ __forceinline__ double src_grav_sink_cool(const DiskParams &P, const DiskScalars &sc, const double4 rot,
                                                     int i, int k, double r, double dV, double zc, double cphi,
                                                     double sphi, const double *pr, double *u, double dt, double *md0,
                                                     double *md1, double *cool) {
  const double rho = pr[RHO], Pp = pr[PPP];
  double Fp;
  add_geom_grav_src(u, pr, r, dV, dt, cphi, sphi, zc, rot.x, rot.y, rot.z, P, sc, &Fp);
  if ((FEAT & F_SINK) && P.sink == 1) {  // get_rho_sink, :15-47
    double xpos = r * cphi, ypos = r * sphi;
    double xbh0 = sc.xbh[0], ybh0 = sc.ybh[0];
    double xbh1 = sc.xbh[1], ybh1 = sc.ybh[1];
    double r0 = sqrt((xpos - xbh0) * (xpos - xbh0) + (ypos - ybh0) * (ypos - ybh0));
    double r1 = sqrt((xpos - xbh1) * (xpos - xbh1) + (ypos - ybh1) * (ypos - ybh1));
    double M0 = sc.gm[0].M, M1 = sc.gm[1].M;
    double Mtotal = M0 + M1;
    double alpha = P.visc;
    double one_o_nu = 1.0 / (alpha * Pp / rho) * sqrt(pow(r0, -3) * M0 / Mtotal + pow(r1, -3) * M1 / Mtotal);
    double t_visc0 = 1.0 * 2. / 3. * r0 * r0 * one_o_nu;
    double t_visc1 = 1.0 * 2. / 3. * r1 * r1 * one_o_nu;
    if (t_visc0 < 10. * dt) t_visc0 = 10. * dt;
    if (t_visc1 < 10. * dt) t_visc1 = 10. * dt;
    int imin = P.NgMinR == 1 ? 0 : P.NgMinR;
    bool interior = i >= imin && i < P.NR - P.NgMaxR && k >= P.NgMinZ && k < P.NZ - P.NgMaxZ;
    for (int pm = 0; pm < P.grav_type && pm < 2; ++pm) {
      double drho = 0.0;
      if (pm == 0 && r0 < 0.5) drho = rho / t_visc0;
      if (pm == 1 && r1 < 0.5) drho = rho / t_visc1;
      u[RHO] -= drho * dt * dV;
      if (interior) {
        if (pm == 0)
          *md0 += drho * dV;
        else
          *md1 += drho * dV;
      }
    }
  }
  if ((FEAT & F_COOL) && P.cooling == 1) {
    if (P.initial_data_type == 0) {
      double cooling_term = (Pp - .01) / (P.gamma - 1.) * dV / 1.e-2;
      if (cooling_term < 0.0) cooling_term = 0.0;
      u[TAU] -= cooling_term * dt;
      if (cool) *cool = cooling_term / dV;
    } else {
      double cooling_cap = u[TAU] * 1.0;
      double cooling_term = 9. / 4. * P.visc * pow(P.porho_r1, -3) / rho * pow(Pp / rho, 4.) * dt * dV;
      if (cooling_term > cooling_cap) cooling_term = cooling_cap;
      u[TAU] -= cooling_term;
      if (cool) *cool = cooling_term / dt / dV;
    }
  }
  return Fp;
}

// cell_add_visc_src for one cell (src/Cell/cell_source.c:216-277): the
// viscous heating of the frame's shear; grad_upp and gradp_urr are the cell's
// r slope of omega and phi slope of v_r
template <int FEAT>
__device__ __forceinline__ void src_visc(const DiskParams &P, const DiskScalars &sc, const double4 rot, double r,
                                         double dV, double cphi, double sphi, const double *pr, double grad_upp,
                                         double gradp_urr, double *u, double dt) {
  const double Pp = pr[PPP], rho = pr[RHO];
  double xpos = r * cphi, ypos = r * sphi;
  double xbh0 = sc.xbh[0], ybh0 = sc.ybh[0];
  double xbh1 = sc.xbh[1], ybh1 = sc.ybh[1];
  double dist_bh0 = sqrt((xpos - xbh0) * (xpos - xbh0) + (ypos - ybh0) * (ypos - ybh0));
  double dist_bh1 = sqrt((xpos - xbh1) * (xpos - xbh1) + (ypos - ybh1) * (ypos - ybh1));
  double eps = P.g_eps;
  double rdrOm_a = rot.y;
  double Sigma_nu;
  if (P.visc_const == 1)
    Sigma_nu = rho * P.visc;
  else
    Sigma_nu = P.visc * Pp /
               sqrt(pow_m1p5(dist_bh0 * dist_bh0 + eps * eps) * sc.gm[0].M +
                    pow_m1p5(dist_bh1 * dist_bh1 + eps * eps) * sc.gm[1].M);
  u[TAU] += dt * dV * Sigma_nu * rdrOm_a * (rdrOm_a + r * grad_upp + gradp_urr);
}

// cell_add_split_fictitious for one cell (cell_add_split_fictitious.c:11-44):
// the exact rotation of (LLL, SRR) by the epicyclic frequency over dt; a no-op
// without a frame rotation (kappa^2 = 0)
template <int FEAT>
__device__ __forceinline__ void src_fict(const double4 rot, double r, double dt, double *u) {
  if (!(FEAT & F_FRAME)) return;
  double rOm = rot.x;
  double Omega = rOm / r;
  double dr_r2Om = 2. * rOm + r * rot.y;
  double kappa2 = 2. * Omega / r * dr_r2Om;
  if (kappa2 > 0.) {
    double kappa = sqrt(kappa2);
    double l0 = u[LLL];
    double rhov0 = u[SRR];
    double skt, ckt;
    sincos(kappa * dt, &skt, &ckt);
    double l = l0 * ckt - u[SRR] * dr_r2Om / kappa * skt;
    double rhov = rhov0 * ckt + 2.0 * Omega * u[LLL] / r / kappa * skt;
    u[LLL] = l;
    u[SRR] = rhov;
  }
}

// The initial-data plugin's single-cell values (src/Cell/initial_data/*.c,
// cell_single_init_*) at a cell of annulus row R with edge tiph and width
// dphi: the shear's from its angle, the others from the row the host
// tabulated (INIT_SET: the plugin writes them).
template <int NQ>
__device__ __forceinline__ void single_init(int type, const InitRow &R, double r, double tiph, double dphi,
                                            double *pr) {
  if (type == ID_SHEAR) {  // cell_init_shear.c:11-55
    const double t = tiph - .5 * dphi;
    const double x = r * cos(t) - 3.;
    const double y = r * sin(t);
    double vy = 0.0;
    if (fabs(y) < 50.) vy = R.c0 * exp(-x * x / R.c1);
    pr[RHO] = 1.0;
    pr[PPP] = 0.01;
    pr[URR] = vy * sin(t);
    pr[UPP] = vy * cos(t) / r - R.a;
    pr[UZZ] = 0.0;
    return;
  }
  if (!(R.flags & INIT_SET)) return;  // the vortex outside its core writes nothing
#pragma unroll
  for (int q = 0; q < 5; ++q) pr[q] = R.prim[q];
  if (type == ID_VORTEX && NQ > 5) pr[5] = cos(tiph - .5 * dphi) < 0.0 ? 0.0 : 1.0;  // cell_init_vortex.c:146,152
}
