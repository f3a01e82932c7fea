/*
 * CPU restatement of the Disco2 hydro substep -- TEST INFRASTRUCTURE ONLY.
 * The product never links or loads this file; it is the checker.
 *
 * Every routine names the reference routine it restates (file:line, paths
 * relative to the reference's src/).  Loop order and floating-point
 * expression association follow the reference exactly, so results are
 * bitwise equal to the reference build on the same host libm.
 *
 * Pinning (see tests/test_oracle_golden.py): every function here is checked
 * against fixtures written by the reference build (tests/golden/ fixtures):
 * geometry, faces, PLM slopes, one substep and whole RK2 steps.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "disco_oracle.h"

#define TWO_PI (2.0 * M_PI) /* the reference's PHIMAX, src/Headers/header.h:21 */
enum { RHO, PPP, URR, UPP, UZZ };
enum { DDD, TAU, SRR, LLL, SZZ };
enum { C_FIXED, C_WCELL, C_RIGID, C_KEPLER, C_OMEGA20, C_MILOS, C_POWERLAW };
enum { A_FIXED, A_KEPLER, A_OMEGA20, A_MILOS, A_POWERLAW };
enum { LEFT, LEFTSTAR, RIGHTSTAR, RIGHT };
enum { BOUND_FIXED, BOUND_OUTFLOW, BOUND_PERIODIC };
enum { SHEAR = 0 };
enum { HLLC, HLL };
#define W_A_MILOS_INDEX 8 /* src/Headers/header.h:25 */
#define TVISC_FAC 1.0     /* src/Headers/header.h:20 */

typedef struct {
  int L, R;
  double r, dphi, dA, deltaL, deltaR, cm;
} OFace;

typedef struct {
  double r, phi, M, omega, RK_r, RK_phi, RK_M, RK_omega, Mdot, Macc;
} OGrav;

struct OracleSim {
  int NR, NZ, NQ, NC, NL, Nz;
  int *N_p, *aoff;
  double *r_faces, *z_faces;
  int ip[OI_COUNT];
  double dp[OD_COUNT];
  double *prim, *cons, *RKcons, *grad, *gradp;
  double *tiph, *RKtiph, *dphi, *wiph, *Cool;
  OGrav gm[2];
  double t, dt, RK;
  OFace *faces[2];
  int nfaces[2];
  int *nri, *nzk;
  int err;
  double torque; /* sum of dt*dV*rho*F_phi*r over interior cells and substeps (the adjoint's loss) */
  /* ledger of the last substep: what the r faces between the inner ghost
   * annulus and the first interior one ([0]) and between the last interior
   * annulus and the outer ghosts ([1]) added to their R cell, dt*dA*F[q]
   * (q < NQ) and dt*dA*Fvisc[SRR, LLL, TAU]; and that substep's torque term */
  double led[2][8 + 3];
  double led_torque;
};

/* sim_FacePos (src/Sim/sim_access.c:26-36): index -1 is the first face */
static inline double rpos(const OracleSim *s, int i) { return s->r_faces[i + 1]; }
static inline double zpos(const OracleSim *s, int k) { return s->z_faces[k + 1]; }
static inline int zid(const OracleSim *s, int k, int i, int j) { return k * s->NL + s->aoff[i] + j; }
#define P(s, q, z) ((s)->prim[(size_t)(q) * (s)->Nz + (z)])
#define U(s, q, z) ((s)->cons[(size_t)(q) * (s)->Nz + (z)])
#define RU(s, q, z) ((s)->RKcons[(size_t)(q) * (s)->Nz + (z)])
#define G(s, q, z) ((s)->grad[(size_t)(q) * (s)->Nz + (z)])
#define GP(s, q, z) ((s)->gradp[(size_t)(q) * (s)->Nz + (z)])

OracleSim *oracle_create(int NR, int NZ, int NQ, const int *N_p, const double *r_faces,
                         const double *z_faces, const int *iparams, const double *dparams) {
  OracleSim *s = calloc(1, sizeof(OracleSim));
  s->NR = NR;
  s->NZ = NZ;
  s->NQ = NQ;
  memcpy(s->ip, iparams, sizeof(s->ip));
  memcpy(s->dp, dparams, sizeof(s->dp));
  s->NC = s->ip[OI_NUM_C];
  s->N_p = malloc(sizeof(int) * NR);
  s->aoff = malloc(sizeof(int) * (NR + 1));
  memcpy(s->N_p, N_p, sizeof(int) * NR);
  s->aoff[0] = 0;
  for (int i = 0; i < NR; ++i) s->aoff[i + 1] = s->aoff[i] + N_p[i];
  s->NL = s->aoff[NR];
  s->Nz = s->NL * NZ;
  s->r_faces = malloc(sizeof(double) * (NR + 1));
  s->z_faces = malloc(sizeof(double) * (NZ + 1));
  memcpy(s->r_faces, r_faces, sizeof(double) * (NR + 1));
  memcpy(s->z_faces, z_faces, sizeof(double) * (NZ + 1));
  size_t nq = (size_t)NQ * s->Nz;
  s->prim = calloc(nq, sizeof(double));
  s->cons = calloc(nq, sizeof(double));
  s->RKcons = calloc(nq, sizeof(double));
  s->grad = calloc(nq, sizeof(double));
  s->gradp = calloc(nq, sizeof(double));
  s->tiph = calloc(s->Nz, sizeof(double));
  s->RKtiph = calloc(s->Nz, sizeof(double));
  s->dphi = calloc(s->Nz, sizeof(double));
  s->wiph = calloc(s->Nz, sizeof(double));
  s->Cool = calloc(s->Nz, sizeof(double));
  s->nri = calloc(NR, sizeof(int));
  s->nzk = calloc(NZ, sizeof(int));
  /* dphi as set by cell_create (src/Cell/cell_create_destroy.c:69,90) */
  for (int k = 0; k < NZ; ++k)
    for (int i = 0; i < NR; ++i) {
      double d = TWO_PI / (double)N_p[i];
      for (int j = 0; j < N_p[i]; ++j) s->dphi[zid(s, k, i, j)] = d;
    }
  return s;
}

void oracle_destroy(OracleSim *s) {
  if (!s) return;
  free(s->N_p); free(s->aoff); free(s->r_faces); free(s->z_faces);
  free(s->prim); free(s->cons); free(s->RKcons); free(s->grad); free(s->gradp);
  free(s->tiph); free(s->RKtiph); free(s->dphi); free(s->wiph); free(s->Cool);
  free(s->nri); free(s->nzk); free(s->faces[0]); free(s->faces[1]);
  free(s);
}

int oracle_nzones(OracleSim *s) { return s->Nz; }
int oracle_last_error(OracleSim *s) { return s->err; }
void oracle_set_tiph(OracleSim *s, const double *t) { memcpy(s->tiph, t, sizeof(double) * s->Nz); }
void oracle_set_dphi(OracleSim *s, const double *d) { memcpy(s->dphi, d, sizeof(double) * s->Nz); }
void oracle_set_prim(OracleSim *s, const double *p) { memcpy(s->prim, p, sizeof(double) * s->NQ * s->Nz); }

static double *field_ptr(OracleSim *s, int f, size_t *n) {
  size_t nq = (size_t)s->NQ * s->Nz;
  switch (f) {
    case 0: *n = nq; return s->prim;
    case 1: *n = nq; return s->cons;
    case 2: *n = nq; return s->RKcons;
    case 3: *n = nq; return s->grad;
    case 4: *n = nq; return s->gradp;
    case 5: *n = s->Nz; return s->tiph;
    case 6: *n = s->Nz; return s->RKtiph;
    case 7: *n = s->Nz; return s->dphi;
    case 8: *n = s->Nz; return s->wiph;
  }
  *n = 0;
  return NULL;
}
void oracle_get(OracleSim *s, int f, double *out) {
  size_t n;
  double *p = field_ptr(s, f, &n);
  if (p) memcpy(out, p, n * sizeof(double));
}
void oracle_put(OracleSim *s, int f, const double *in) {
  size_t n;
  double *p = field_ptr(s, f, &n);
  if (p) memcpy(p, in, n * sizeof(double));
}
void oracle_scalars(OracleSim *s, double *o) {
  o[0] = s->t; o[1] = s->dt; o[2] = s->RK;
  for (int p = 0; p < 2; ++p) {
    double *g = o + 3 + 10 * p;
    g[0] = s->gm[p].r; g[1] = s->gm[p].phi; g[2] = s->gm[p].M; g[3] = s->gm[p].omega;
    g[4] = s->gm[p].RK_r; g[5] = s->gm[p].RK_phi; g[6] = s->gm[p].RK_M; g[7] = s->gm[p].RK_omega;
    g[8] = s->gm[p].Mdot; g[9] = s->gm[p].Macc;
  }
}
void oracle_set_scalars(OracleSim *s, const double *o) {
  s->t = o[0]; s->dt = o[1]; s->RK = o[2];
  for (int p = 0; p < 2; ++p) {
    const double *g = o + 3 + 10 * p;
    s->gm[p].r = g[0]; s->gm[p].phi = g[1]; s->gm[p].M = g[2]; s->gm[p].omega = g[3];
    s->gm[p].RK_r = g[4]; s->gm[p].RK_phi = g[5]; s->gm[p].RK_M = g[6]; s->gm[p].RK_omega = g[7];
    s->gm[p].Mdot = g[8]; s->gm[p].Macc = g[9];
  }
}

/* ---------------- frame rotation: src/Sim/sim_w_a.c:8-73 ---------------- */
static double rOm_a(const OracleSim *s, double r, double a) {
  switch (s->ip[OI_W_ANALYTIC_TYPE]) {
    case A_FIXED: return 0.0;
    case A_KEPLER: return pow(r, -0.5);
    case A_OMEGA20: return r * 20;
    case A_MILOS:
      return r * (pow(r, W_A_MILOS_INDEX - 1.5) + pow(a, W_A_MILOS_INDEX - 1.5)) /
             (pow(r, W_A_MILOS_INDEX) + pow(a, W_A_MILOS_INDEX));
    case A_POWERLAW: return pow(r, -1.0);
  }
  return 0.0;
}
static double rdrOm_a(const OracleSim *s, double r, double a) {
  switch (s->ip[OI_W_ANALYTIC_TYPE]) {
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
  return 0.0;
}
static double dtOm_a(const OracleSim *s, double r, double a) {
  if (s->ip[OI_W_ANALYTIC_TYPE] == A_MILOS) {
    double a0 = 1.0;
    double tau = s->dp[OD_OrbShrinkTscale];
    double dta_o_a = -.25 * pow(a / a0, -4) / tau;
    int n = W_A_MILOS_INDEX;
    double Omega_a = (pow(r, n - 1.5) + pow(a, n - 1.5)) / (pow(r, n) + pow(a, n));
    double arn = pow(a / r, n);
    if (a > 0.0) return dta_o_a * (Omega_a * (-1.5 + n / (1. + arn)) + (1.5 - n) * pow(r, -1.5) / (1. + arn));
    return 0.0;
  }
  return 0.0;
}

/* ---------------- GravMass: src/GravMass/ ---------------- */
static void grav_init(OracleSim *s) { /* GravMass_init.c:22-84 */
  int ty = s->ip[OI_GravMassType];
  memset(s->gm, 0, sizeof(s->gm));
  if (ty == 1) {
    s->gm[0].M = 1.0; s->gm[0].r = 0.0; s->gm[0].phi = 0.0; s->gm[0].omega = 0.0;
    s->gm[1].M = 0.0; s->gm[1].r = 1.0; s->gm[1].phi = 0.0; s->gm[1].omega = 0.0;
  } else if (ty == 2) {
    double Mtotal = 1.0, sep = 1.0;
    double massratio = s->dp[OD_MassRatio];
    double M0 = Mtotal / (1. + massratio);
    double M1 = Mtotal / (1. + 1. / massratio);
    double r0 = M1 / Mtotal * sep;
    double r1 = M0 / Mtotal * sep;
    double om2 = 1. / sep / sep / sep;
    s->gm[0].M = M0; s->gm[0].r = r0; s->gm[0].phi = 0.0; s->gm[0].omega = sqrt(om2);
    s->gm[1].M = M1; s->gm[1].r = r1; s->gm[1].phi = M_PI; s->gm[1].omega = sqrt(om2);
  }
}
static void grav_clean_pi(OracleSim *s) { /* GravMass_misc.c:9-17 */
  for (int p = 0; p < 2; ++p) {
    double phi = s->gm[p].phi;
    while (phi > 2. * M_PI) phi -= 2. * M_PI;
    while (phi < 0.0) phi += 2. * M_PI;
    s->gm[p].phi = phi;
  }
}
static void grav_copy(OracleSim *s) { /* GravMass_misc.c:20-28 */
  for (int p = 0; p < 2; ++p) {
    s->gm[p].RK_r = s->gm[p].r; s->gm[p].RK_phi = s->gm[p].phi;
    s->gm[p].RK_M = s->gm[p].M; s->gm[p].RK_omega = s->gm[p].omega;
  }
}
static void grav_move(OracleSim *s, double t, double dt) { /* GravMass_misc.c:31-64 */
  double m0 = s->gm[0].M, m1 = s->gm[1].M;
  double M = m0 + m1;
  double a_0 = 1.0;
  double tau_0 = s->dp[OD_OrbShrinkTscale];
  double T0 = s->dp[OD_OrbShrinkT0];
  double a, omega;
  if ((t - T0) / tau_0 < 1.) {
    a = a_0 * pow((1.0 - (t - T0) / tau_0), 0.25);
    omega = pow(a / M, -1.5);
  } else {
    a = 0.0;
    omega = 0.0;
  }
  double r0, r1;
  if (M < 1.e-12) {
    r0 = 0.0;
    r1 = 0.0;
  } else {
    r0 = m1 / M * a;
    r1 = m0 / M * a;
  }
  s->gm[0].phi += omega * dt;
  s->gm[1].phi += omega * dt;
  s->gm[0].r = r0;
  s->gm[1].r = r1;
  s->gm[0].omega = omega;
  s->gm[1].omega = omega;
}
static void grav_update_RK(OracleSim *s, double RK) { /* GravMass_misc.c:66-77 */
  for (int p = 0; p < 2; ++p) {
    OGrav *g = &s->gm[p];
    while (g->phi - g->RK_phi > M_PI) g->RK_phi += 2. * M_PI;
    while (g->phi - g->RK_phi < -M_PI) g->RK_phi -= 2. * M_PI;
    g->r = (1.0 - RK) * g->r + RK * g->RK_r;
    g->phi = (1.0 - RK) * g->phi + RK * g->RK_phi;
    g->M = (1.0 - RK) * g->M + RK * g->RK_M;
    g->omega = (1.0 - RK) * g->omega + RK * g->RK_omega;
  }
}
static double grav_dist(const OracleSim *s, int p, double r, double phi, double z) { /* GravMass_access.c:25-31 */
  double xbh = s->gm[p].r * cos(s->gm[p].phi);
  double ybh = s->gm[p].r * sin(s->gm[p].phi);
  double x = r * cos(phi);
  double y = r * sin(phi);
  return sqrt((x - xbh) * (x - xbh) + (y - ybh) * (y - ybh) + z * z);
}

/* ---------------- conversions: src/Cell/cell_conversion.c ---------------- */
static void prim2cons(const OracleSim *s, const double *prim, double *cons, double r, double dV) { /* :11-33 */
  double G = s->dp[OD_GAMMALAW];
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
  for (int q = s->NC; q < s->NQ; ++q) cons[q] = prim[q] * cons[DDD];
}
static void cons2prim(const OracleSim *s, const double *cons, double *prim, double r, double dV) { /* :57-103 */
  double CS_FLOOR = s->dp[OD_CS_FLOOR], CS_CAP = s->dp[OD_CS_CAP];
  double RHO_FLOOR = s->dp[OD_RHO_FLOOR], VEL_CAP = s->dp[OD_VEL_CAP];
  double G = s->dp[OD_GAMMALAW];
  double rho = cons[DDD] / dV;
  if (rho < RHO_FLOOR) rho = RHO_FLOOR;
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
  if (Pp < CS_FLOOR * CS_FLOOR * rho / G) Pp = CS_FLOOR * CS_FLOOR * rho / G;
  if (Pp > CS_CAP * CS_CAP * rho / G) Pp = CS_CAP * CS_CAP * rho / G;
  if (v_magnitude > VEL_CAP) {
    vr = vr * VEL_CAP / v_magnitude;
    vp = vp * VEL_CAP / v_magnitude;
    vz = vz * VEL_CAP / v_magnitude;
  }
  prim[RHO] = rho;
  prim[PPP] = Pp;
  prim[URR] = vr;
  prim[UPP] = vp / r;
  prim[UZZ] = vz;
  for (int q = s->NC; q < s->NQ; ++q) prim[q] = cons[q] / cons[DDD];
}

static void gather(const OracleSim *s, const double *f, int z, double *v) {
  for (int q = 0; q < s->NQ; ++q) v[q] = f[(size_t)q * s->Nz + z];
}
static void scatter(OracleSim *s, double *f, int z, const double *v) {
  for (int q = 0; q < s->NQ; ++q) f[(size_t)q * s->Nz + z] = v[q];
}

static void calc_cons(OracleSim *s) { /* cell_conversion.c:35-55 */
  double pr[16], cs[16];
  for (int k = 0; k < s->NZ; ++k) {
    double zp = zpos(s, k), zm = zpos(s, k - 1);
    double dz = zp - zm;
    for (int i = 0; i < s->NR; ++i) {
      double rp = rpos(s, i), rm = rpos(s, i - 1);
      double r = .5 * (rp + rm);
      for (int j = 0; j < s->N_p[i]; ++j) {
        int z = zid(s, k, i, j);
        double dV = .5 * (rp * rp - rm * rm) * s->dphi[z] * dz;
        gather(s, s->prim, z, pr);
        prim2cons(s, pr, cs, r, dV);
        scatter(s, s->cons, z, cs);
      }
    }
  }
}
static void calc_prim(OracleSim *s) { /* cell_conversion.c:105-122 */
  double pr[16], cs[16];
  for (int k = 0; k < s->NZ; ++k) {
    double zm = zpos(s, k - 1), zp = zpos(s, k);
    double dz = zp - zm;
    for (int i = 0; i < s->NR; ++i) {
      double rm = rpos(s, i - 1), rp = rpos(s, i);
      double r = .5 * (rp + rm);
      for (int j = 0; j < s->N_p[i]; ++j) {
        int z = zid(s, k, i, j);
        double dV = .5 * (rp * rp - rm * rm) * s->dphi[z] * dz;
        gather(s, s->cons, z, cs);
        cons2prim(s, cs, pr, r, dV);
        scatter(s, s->prim, z, pr);
      }
    }
  }
}

/* ---------------- per-cell bookkeeping: src/Cell/cell_misc.c ---------------- */
static void clean_pi(OracleSim *s) { /* :12-26 */
  for (int z = 0; z < s->Nz; ++z) {
    double phi = s->tiph[z];
    while (phi > TWO_PI) phi -= TWO_PI;
    while (phi < 0.0) phi += TWO_PI;
    s->tiph[z] = phi;
  }
}
static void cell_copy(OracleSim *s) { /* :28-41 */
  memcpy(s->RKcons, s->cons, sizeof(double) * s->NQ * s->Nz);
  memcpy(s->RKtiph, s->tiph, sizeof(double) * s->Nz);
}
static void adjust_RK_cons(OracleSim *s, double RK) { /* :43-57 */
  size_t n = (size_t)s->NQ * s->Nz;
  for (size_t x = 0; x < n; ++x) s->cons[x] = (1.0 - RK) * s->cons[x] + RK * s->RKcons[x];
}
static void update_phi(OracleSim *s, double RK, double dt) { /* :59-79 */
  double a = s->gm[0].r + s->gm[1].r;
  for (int k = 0; k < s->NZ; ++k)
    for (int i = 0; i < s->NR; ++i) {
      double rm = rpos(s, i - 1), rp = rpos(s, i);
      double r = .5 * (rm + rp);
      for (int j = 0; j < s->N_p[i]; ++j) {
        int z = zid(s, k, i, j);
        while (s->tiph[z] - s->RKtiph[z] > TWO_PI / 2.) s->RKtiph[z] += TWO_PI;
        while (s->tiph[z] - s->RKtiph[z] < -TWO_PI / 2.) s->RKtiph[z] -= TWO_PI;
        s->tiph[z] = (1.0 - RK) * s->tiph[z] + RK * s->RKtiph[z];
        double w_total = s->wiph[z] + rOm_a(s, r, a);
        s->tiph[z] += w_total * dt / r;
      }
    }
}
static void update_dphi(OracleSim *s) { /* :81-96 */
  for (int k = 0; k < s->NZ; ++k)
    for (int i = 0; i < s->NR; ++i)
      for (int j = 0; j < s->N_p[i]; ++j) {
        int jm = j - 1;
        if (jm == -1) jm = s->N_p[i] - 1;
        double d = s->tiph[zid(s, k, i, j)] - s->tiph[zid(s, k, i, jm)];
        while (d > TWO_PI) d -= TWO_PI;
        while (d < 0.0) d += TWO_PI;
        s->dphi[zid(s, k, i, j)] = d;
      }
}

/* src/Cell/cell_set_w.c:12-102 (after cell_clear_w, cell_misc.c:134-143) */
void oracle_set_w(OracleSim *s) {
  memset(s->wiph, 0, sizeof(double) * s->Nz);
  int ty = s->ip[OI_W_NUMERIC_TYPE];
  if (ty == C_WCELL) {
    s->err = 11;
    return;
  }
  for (int k = 0; k < s->NZ; ++k)
    for (int i = 0; i < s->NR; ++i) {
      double rp = rpos(s, i), rm = rpos(s, i - 1);
      double r = 0.5 * (rm + rp);
      double w = 0.0;
      if (ty == C_RIGID) {
        double ww = 0.0, Mring = 0.0;
        for (int j = 0; j < s->N_p[i]; ++j) {
          int z = zid(s, k, i, j);
          double vp = r * P(s, UPP, z);
          double m = U(s, DDD, z);
          Mring += m;
          ww += vp * m;
        }
        w = ww / Mring;
      } else if (ty == C_KEPLER) {
        w = pow(r, -0.5);
      } else if (ty == C_OMEGA20) {
        w = r * 20.;
      } else if (ty == C_MILOS) {
        w = (1. - exp(-pow(r, 1.5))) / sqrt(r);
      } else if (ty == C_POWERLAW) {
        w = pow(r, -1.0);
      } else {
        continue;
      }
      for (int j = 0; j < s->N_p[i]; ++j) s->wiph[zid(s, k, i, j)] = w;
    }
}

/* ---------------- CFL: src/Cell/cell_mindt.c:10-100 ---------------- */
static double maxvel(const OracleSim *s, int z, double w, double r) {
  double Pp = P(s, PPP, z);
  double rho = P(s, RHO, z);
  double vp = P(s, UPP, z) * r - w;
  double vr = P(s, URR, z);
  double vz = P(s, UZZ, z);
  double cs = sqrt(s->dp[OD_GAMMALAW] * Pp / rho);
  double cf2 = cs * cs;
  return sqrt(cf2) + sqrt(vr * vr + vp * vp + vz * vz);
}
double oracle_mindt(OracleSim *s) {
  double dt_m = 1.e100;
  for (int k = s->ip[OI_NgMinZ]; k < s->NZ - s->ip[OI_NgMaxZ]; ++k) {
    double zm = zpos(s, k - 1), zp = zpos(s, k);
    double dz = zp - zm;
    for (int i = s->ip[OI_NgMinR]; i < s->NR - s->ip[OI_NgMaxR]; ++i) {
      double rm = rpos(s, i - 1), rp = rpos(s, i);
      double dr = rp - rm;
      double r = .5 * (rp + rm);
      for (int j = 0; j < s->N_p[i]; ++j) {
        int jm = j - 1;
        if (j == 0) jm = s->N_p[i] - 1;
        int z = zid(s, k, i, j);
        double w = .5 * (s->wiph[z] + s->wiph[zid(s, k, i, jm)]);
        double dx = dr;
        double rdphi = .5 * (rp + rm) * s->dphi[z];
        if (rdphi < dr) dx = rdphi;
        if (dx > dz) dx = dz;
        double a = maxvel(s, z, w, r);
        double rho = P(s, RHO, z);
        double Pp = P(s, PPP, z);
        double dt = s->dp[OD_CFL] * dx / a;
        if (s->dp[OD_EXPLICIT_VISCOSITY] > 0.0) {
          double nu;
          if (s->ip[OI_VISC_CONST] == 1) {
            nu = s->dp[OD_EXPLICIT_VISCOSITY];
          } else {
            double tiph = s->tiph[z] - 0.5 * s->dphi[z];
            if (s->ip[OI_InitialDataType] == SHEAR) {
              nu = s->dp[OD_EXPLICIT_VISCOSITY] * s->dp[OD_GAMMALAW] * Pp / rho * pow(fabs(r * cos(tiph)), 2.0);
              if (r * cos(tiph) > 20.) nu = 0.000000001;
            } else {
              double M0 = s->gm[0].M, M1 = s->gm[1].M;
              double d0 = grav_dist(s, 0, r, tiph, 0.);
              double d1 = grav_dist(s, 1, r, tiph, 0.);
              double alpha = s->dp[OD_EXPLICIT_VISCOSITY];
              double eps = s->dp[OD_G_EPS];
              nu = alpha * Pp / rho / sqrt(pow(d0 * d0 + eps * eps, -1.5) * M0 + pow(d1 * d1 + eps * eps, -1.5) * M1);
            }
          }
          double dt_visc = .25 * dx * dx / nu;
          dt = dt / (1. + dt / dt_visc);
        }
        if (dt_m > dt) dt_m = dt;
      }
    }
  }
  return dt_m;
}

/* ---------------- faces: src/Face/face_create_destroy.c ---------------- */
static void add_face(OFace *f, int n, int L, int R, double r, double deltaL, double deltaR,
                     double dphi, double tp, double deltaPerp) { /* :11-20 */
  f[n].L = L;
  f[n].R = R;
  f[n].r = r;
  f[n].deltaL = deltaL;
  f[n].deltaR = deltaR;
  f[n].dphi = dphi;
  f[n].dA = r * dphi * deltaPerp;
  f[n].cm = tp - .5 * dphi;
}
/* build_jloop, :22-128.  Inner annulus (k,i); outer (k,i+1) for r, (k+1,i) for z. */
static void build_jloop(OracleSim *s, int *pn, int i, int k, int rDir, int zDir, OFace *f, int mode) {
  double deltaL = 0, deltaR = 0, deltaPerp = 0, r = 0;
  if (rDir) {
    deltaL = .5 * (rpos(s, i) - rpos(s, i - 1));
    deltaR = .5 * (rpos(s, i + 1) - rpos(s, i));
    r = rpos(s, i);
    double zp = zpos(s, k), zm = zpos(s, k - 1);
    deltaPerp = zp - zm;
  }
  if (zDir) {
    deltaL = .5 * (zpos(s, k) - zpos(s, k - 1));
    deltaR = .5 * (zpos(s, k + 1) - zpos(s, k));
    double rp = rpos(s, i), rm = rpos(s, i - 1);
    deltaPerp = rp - rm;
    r = .5 * (rp + rm);
  }
  int Nin = s->N_p[i], Nout = s->N_p[i + rDir];
  int zin0 = zid(s, k, i, 0), zout0 = zid(s, k + zDir, i + rDir, 0);
  const double *T = s->tiph, *D = s->dphi;
  double p0 = T[zin0 + Nin - 1];
  int jpmin = 0;
  double dpmin = TWO_PI;
  int jp;
  for (jp = 0; jp < Nout; ++jp) {
    double dp = T[zout0 + jp] - p0;
    while (dp > TWO_PI) dp -= TWO_PI;
    while (dp < 0.0) dp += TWO_PI;
    if (dpmin > dp) {
      dpmin = dp;
      jpmin = jp;
    }
  }
  jp = jpmin;
  for (int j = 0; j < Nin; ++j) {
    int jm = j - 1;
    if (jm < 0) jm = Nin - 1;
    double dp = T[zout0 + jp] - T[zin0 + jm];
    while (dp > TWO_PI) dp -= TWO_PI;
    while (dp < 0.0) dp += TWO_PI;
    if (D[zin0 + j] < dp) {
      if (mode == 1) add_face(f, *pn, zin0 + j, zout0 + jp, r, deltaL, deltaR, D[zin0 + j], T[zin0 + j], deltaPerp);
      ++(*pn);
    } else {
      if (mode == 1) add_face(f, *pn, zin0 + j, zout0 + jp, r, deltaL, deltaR, dp, T[zout0 + jp], deltaPerp);
      ++(*pn);
      ++jp;
      if (jp == Nout) jp = 0;
      dp = T[zin0 + j] - T[zout0 + jp];
      while (dp > TWO_PI / 2.) dp -= TWO_PI;
      while (dp < -TWO_PI / 2.) dp += TWO_PI;
      while (dp > 0.0) {
        if (mode == 1) add_face(f, *pn, zin0 + j, zout0 + jp, r, deltaL, deltaR, D[zout0 + jp], T[zout0 + jp], deltaPerp);
        ++(*pn);
        ++jp;
        if (jp == Nout) jp = 0;
        dp = T[zin0 + j] - T[zout0 + jp];
        while (dp > TWO_PI / 2.) dp -= TWO_PI;
        while (dp < -TWO_PI / 2.) dp += TWO_PI;
      }
      dp = D[zout0 + jp] + dp;
      if (mode == 1) add_face(f, *pn, zin0 + j, zout0 + jp, r, deltaL, deltaR, dp, T[zin0 + j], deltaPerp);
      ++(*pn);
    }
  }
}
int oracle_face_create(OracleSim *s, int dir) { /* face_create, :130-193 */
  int n = 0;
  if (dir == 0) {
    for (int i = 0; i < s->NR - 1; ++i) {
      s->nri[i] = n;
      for (int k = 0; k < s->NZ; ++k) build_jloop(s, &n, i, k, 1, 0, NULL, 0);
    }
    s->nri[s->NR - 1] = n;
    free(s->faces[0]);
    s->faces[0] = malloc(sizeof(OFace) * (n + 1));
    n = 0;
    for (int i = 0; i < s->NR - 1; ++i)
      for (int k = 0; k < s->NZ; ++k) build_jloop(s, &n, i, k, 1, 0, s->faces[0], 1);
    s->nfaces[0] = n;
  } else {
    if (s->ip[OI_NglobalZ] > 1) {
      for (int k = 0; k < s->NZ - 1; ++k) {
        s->nzk[k] = n;
        for (int i = 0; i < s->NR; ++i) build_jloop(s, &n, i, k, 0, 1, NULL, 0);
      }
      s->nzk[s->NZ - 1] = n;
      free(s->faces[1]);
      s->faces[1] = malloc(sizeof(OFace) * (n + 1));
      n = 0;
      for (int k = 0; k < s->NZ - 1; ++k)
        for (int i = 0; i < s->NR; ++i) build_jloop(s, &n, i, k, 0, 1, s->faces[1], 1);
    }
    s->nfaces[1] = n;
  }
  return n;
}
static void face_export(OracleSim *s, int dir, int *LR, double *geo, int *offsets) {
  OFace *f = s->faces[dir];
  for (int n = 0; n < s->nfaces[dir]; ++n) {
    LR[2 * n] = f[n].L;
    LR[2 * n + 1] = f[n].R;
    geo[6 * n + 0] = f[n].r;
    geo[6 * n + 1] = f[n].dphi;
    geo[6 * n + 2] = f[n].dA;
    geo[6 * n + 3] = f[n].deltaL;
    geo[6 * n + 4] = f[n].deltaR;
    geo[6 * n + 5] = f[n].cm;
  }
  if (dir == 0)
    memcpy(offsets, s->nri, sizeof(int) * s->NR);
  else
    memcpy(offsets, s->nzk, sizeof(int) * s->NZ);
}
void oracle_face_export(OracleSim *s, int dir, int *LR, double *geo, int *offsets) {
  face_export(s, dir, LR, geo, offsets);
}

/* ---------------- PLM: src/Cell/cell_plm.c ---------------- */
void oracle_plm_p(OracleSim *s) { /* :116-161 */
  double PLM = s->dp[OD_PLM];
  for (int k = 0; k < s->NZ; ++k)
    for (int i = 0; i < s->NR; ++i) {
      int N = s->N_p[i];
      for (int j = 0; j < N; ++j) {
        int c = zid(s, k, i, j);
        int cL, cR;
        if (j == 0) {
          cL = zid(s, k, i, N - 1);
          cR = zid(s, k, i, j + 1);
        } else if (j == N - 1) {
          cL = zid(s, k, i, j - 1);
          cR = zid(s, k, i, 0);
        } else {
          cL = zid(s, k, i, j - 1);
          cR = zid(s, k, i, j + 1);
        }
        double dpL = .5 * (s->dphi[c] + s->dphi[cL]);
        double dpR = .5 * (s->dphi[c] + s->dphi[cR]);
        for (int q = 0; q < s->NQ; ++q) {
          double sL = (P(s, q, c) - P(s, q, cL)) / dpL;
          double sR = (P(s, q, cR) - P(s, q, c)) / dpR;
          double sM = (P(s, q, cR) - P(s, q, cL)) / (dpL + dpR);
          double sl = PLM * sL;
          if (fabs(sl) > PLM * fabs(sR)) sl = PLM * sR;
          if (fabs(sl) > fabs(sM)) sl = sM;
          if (sL * sR < 0.0) sl = 0.0;
          GP(s, q, c) = sl;
        }
      }
    }
}
static inline double wrap_pm(double d) {
  while (d > TWO_PI / 2.) d -= TWO_PI;
  while (d < -TWO_PI / 2.) d += TWO_PI;
  return d;
}
void oracle_plm_rz(OracleSim *s, int dir) { /* :13-114 */
  double PLM = s->dp[OD_PLM];
  OFace *f = s->faces[dir];
  int Nf = s->nfaces[dir];
  memset(s->grad, 0, sizeof(double) * s->NQ * s->Nz);
  for (int n = 0; n < Nf; ++n) {
    int cL = f[n].L, cR = f[n].R;
    double pL = s->tiph[cL] - .5 * s->dphi[cL];
    double pR = s->tiph[cR] - .5 * s->dphi[cR];
    double dpL = wrap_pm(f[n].cm - pL);
    double dpR = wrap_pm(pR - f[n].cm);
    double dA = f[n].dA;
    for (int q = 0; q < s->NQ; ++q) {
      double WL = P(s, q, cL) + dpL * GP(s, q, cL);
      double WR = P(s, q, cR) - dpR * GP(s, q, cR);
      double S = (WR - WL) / (f[n].deltaR + f[n].deltaL);
      G(s, q, cL) += S * dA;
      G(s, q, cR) += S * dA;
    }
  }
  for (int k = 0; k < s->NZ; ++k) {
    double zm = zpos(s, k - 1), zp = zpos(s, k);
    double dz = zp - zm;
    for (int i = 0; i < s->NR; ++i) {
      double rm = rpos(s, i - 1), rp = rpos(s, i);
      for (int j = 0; j < s->N_p[i]; ++j) {
        int z = zid(s, k, i, j);
        double dp = s->dphi[z];
        double dAtot;
        if (dir == 0) {
          if (i == 0)
            dAtot = dz * rp * dp; /* single rank: always the inner r boundary */
          else if (i == s->NR - 1)
            dAtot = dz * rm * dp;
          else
            dAtot = dz * (rp + rm) * dp;
        } else {
          dAtot = 2. * .5 * (rp * rp - rm * rm) * dp;
        }
        for (int q = 0; q < s->NQ; ++q) G(s, q, z) /= dAtot;
      }
    }
  }
  for (int n = 0; n < Nf; ++n) {
    int cL = f[n].L, cR = f[n].R;
    double pL = s->tiph[cL] - .5 * s->dphi[cL];
    double pR = s->tiph[cR] - .5 * s->dphi[cR];
    double dpL = wrap_pm(f[n].cm - pL);
    double dpR = wrap_pm(pR - f[n].cm);
    for (int q = 0; q < s->NQ; ++q) {
      double WL = P(s, q, cL) + dpL * GP(s, q, cL);
      double WR = P(s, q, cR) - dpR * GP(s, q, cR);
      double S = (WR - WL) / (f[n].deltaR + f[n].deltaL);
      double SL = G(s, q, cL);
      double SR = G(s, q, cR);
      if (S * SL < 0.0)
        G(s, q, cL) = 0.0;
      else if (fabs(PLM * S) < fabs(SL))
        G(s, q, cL) = PLM * S;
      if (S * SR < 0.0)
        G(s, q, cR) = 0.0;
      else if (fabs(PLM * S) < fabs(SR))
        G(s, q, cR) = PLM * S;
    }
  }
}

/* ---------------- Riemann solver: src/Riemann/ ---------------- */
typedef struct {
  int cL, cR;
  double r, r_cell_L, r_cell_R, dA, cm;
  double primL[16], primR[16], UL[16], UR[16], Ustar[16], FL[16], FR[16], Fstar[16], F[16], Fvisc[16];
  double Sl, Sr, Ss;
  int state;
  int n[3];
} ORiemann;

static void riemann_init(ORiemann *R) { memset(R, 0, sizeof(*R)); } /* Riemann_create_destroy.c:9-39 */

static void setup_rz(OracleSim *s, ORiemann *R, int dir, int n, int direction) { /* Riemann_setup.c:11-44 */
  OFace *f = &s->faces[dir][n];
  R->n[direction] = 1;
  double deltaL = f->deltaL, deltaR = f->deltaR;
  R->cL = f->L;
  R->cR = f->R;
  double pL = s->tiph[R->cL] - .5 * s->dphi[R->cL];
  double pR = s->tiph[R->cR] - .5 * s->dphi[R->cR];
  double dpL = f->cm - pL;
  double dpR = -f->cm + pR;
  dpL = wrap_pm(dpL);
  dpR = wrap_pm(dpR);
  R->r = f->r;
  R->dA = f->dA;
  R->cm = f->cm;
  if (direction == 0) {
    R->r_cell_L = R->r - deltaL;
    R->r_cell_R = R->r + deltaR;
  } else {
    R->r_cell_L = R->r;
    R->r_cell_R = R->r;
  }
  for (int q = 0; q < s->NQ; ++q) {
    R->primL[q] = P(s, q, R->cL) + G(s, q, R->cL) * deltaL + GP(s, q, R->cL) * dpL;
    R->primR[q] = P(s, q, R->cR) - G(s, q, R->cR) * deltaR - GP(s, q, R->cR) * dpR;
  }
}
static void setup_p(OracleSim *s, ORiemann *R, int i, int j_low, int k) { /* Riemann_setup.c:46-80 */
  R->n[1] = 1;
  int j_hi = (j_low == s->N_p[i] - 1) ? 0 : j_low + 1;
  R->cL = zid(s, k, i, j_low);
  R->cR = zid(s, k, i, j_hi);
  double dpL = s->dphi[R->cL];
  double dpR = s->dphi[R->cR];
  double zm = zpos(s, k - 1), zp = zpos(s, k);
  double dz = zp - zm;
  double rm = rpos(s, i - 1), rp = rpos(s, i);
  double dr = rp - rm;
  double r = .5 * (rp + rm);
  R->dA = dr * dz;
  R->r = r;
  R->cm = s->tiph[R->cL];
  R->r_cell_L = r;
  R->r_cell_R = r;
  for (int q = 0; q < s->NQ; ++q) {
    R->primL[q] = P(s, q, R->cL) + 0.5 * GP(s, q, R->cL) * dpL;
    R->primR[q] = P(s, q, R->cR) - 0.5 * GP(s, q, R->cR) * dpR;
  }
}
static void LR_speed(const double *prim, double r, const int *n, double G, double *p_vn, double *p_cf2) { /* Riemann_routines.c:14-27 */
  double Pp = prim[PPP];
  double rho = prim[RHO];
  double vr = prim[URR];
  double vp = r * prim[UPP];
  double vz = prim[UZZ];
  double vn = vr * n[0] + vp * n[1] + vz * n[2];
  double cf2 = G * (Pp / rho);
  *p_vn = vn;
  *p_cf2 = cf2;
}
static void set_vel(ORiemann *R, double r, double G) { /* Riemann_routines.c:30-55 */
  double Sl, Sr, Ss;
  double vnL, cf2L, vnR, cf2R;
  LR_speed(R->primL, r, R->n, G, &vnL, &cf2L);
  Sl = vnL - sqrt(cf2L);
  Sr = vnL + sqrt(cf2L);
  LR_speed(R->primR, r, R->n, G, &vnR, &cf2R);
  if (Sl > vnR - sqrt(cf2R)) Sl = vnR - sqrt(cf2R);
  if (Sr < vnR + sqrt(cf2R)) Sr = vnR + sqrt(cf2R);
  Ss = (R->primR[RHO] * vnR * (Sr - vnR) - R->primL[RHO] * vnL * (Sl - vnL) + R->primL[PPP] - R->primR[PPP]) /
       (R->primR[RHO] * (Sr - vnR) - R->primL[RHO] * (Sl - vnL));
  R->Sl = Sl;
  R->Sr = Sr;
  R->Ss = Ss;
}
static void set_state(ORiemann *R, double w) { /* Riemann_routines.c:59-71 */
  if (w < R->Sl)
    R->state = LEFT;
  else if (w > R->Sr)
    R->state = RIGHT;
  else if (w < R->Ss)
    R->state = LEFTSTAR;
  else
    R->state = RIGHTSTAR;
}
static void star_hll(const OracleSim *s, ORiemann *R) { /* Riemann_routines.c:74-85 */
  double Sr = R->Sr, Sl = R->Sl;
  for (int q = 0; q < s->NQ; ++q) {
    R->Ustar[q] = (Sr * R->UR[q] - Sl * R->UL[q] + R->FL[q] - R->FR[q]) / (Sr - Sl);
    R->Fstar[q] = (Sr * R->FL[q] - Sl * R->FR[q] - Sr * Sl * (R->UL[q] - R->UR[q])) / (Sr - Sl);
  }
}
static void star_hllc(const OracleSim *s, ORiemann *R, double G) { /* Riemann_routines.c:87-154 */
  double r = R->r;
  const double *prim;
  double Sk;
  const double *Uk, *Fk;
  if (R->state == LEFTSTAR) {
    prim = R->primL; Sk = R->Sl; Uk = R->UL; Fk = R->FL;
  } else {
    prim = R->primR; Sk = R->Sr; Uk = R->UR; Fk = R->FR;
  }
  double Ss = R->Ss;
  double rho = prim[RHO];
  double vr = prim[URR];
  double vp = prim[UPP] * r;
  double vz = prim[UZZ];
  double Pp = prim[PPP];
  double v2 = vr * vr + vp * vp + vz * vz;
  double vn = vr * R->n[0] + vp * R->n[1] + vz * R->n[2];
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
  double mn = Msr * R->n[0] + Msp * R->n[1] + Msz * R->n[2];
  Msr += R->n[0] * (Msn - mn);
  Msp += R->n[1] * (Msn - mn);
  Msz += R->n[2] * (Msn - mn);
  R->Ustar[DDD] = Dstar;
  R->Ustar[SRR] = Msr;
  R->Ustar[LLL] = r * Msp;
  R->Ustar[SZZ] = Msz;
  R->Ustar[TAU] = Estar;
  for (int q = s->NC; q < s->NQ; ++q) R->Ustar[q] = prim[q] * R->Ustar[DDD];
  for (int q = 0; q < s->NQ; ++q) R->Fstar[q] = Fk[q] + Sk * (R->Ustar[q] - Uk[q]);
}
static void set_flux(const OracleSim *s, ORiemann *R, double G, int side) { /* Riemann_routines.c:157-191 */
  double r = R->r;
  const double *prim = side == LEFT ? R->primL : R->primR;
  double *F = side == LEFT ? R->FL : R->FR;
  double rho = prim[RHO];
  double Pp = prim[PPP];
  double vr = prim[URR];
  double vp = prim[UPP] * r;
  double vz = prim[UZZ];
  double vn = vr * R->n[0] + vp * R->n[1] + vz * R->n[2];
  double rhoe = Pp / (G - 1.);
  double v2 = vr * vr + vp * vp + vz * vz;
  F[DDD] = rho * vn;
  F[SRR] = rho * vr * vn + Pp * R->n[0];
  F[LLL] = r * (rho * vp * vn + Pp * R->n[1]);
  F[SZZ] = rho * vz * vn + Pp * R->n[2];
  F[TAU] = (.5 * rho * v2 + rhoe + Pp) * vn;
  for (int q = s->NC; q < s->NQ; ++q) F[q] = prim[q] * F[DDD];
}
static int visc_flux(OracleSim *s, ORiemann *R) { /* Riemann_routines.c:193-311 */
  int NQ = s->NQ;
  double r = R->r;
  double tiph = R->cm;
  double M0 = s->gm[0].M, M1 = s->gm[1].M;
  double a = s->gm[0].r + s->gm[1].r;
  double dist_bh0 = grav_dist(s, 0, r, tiph, 0.);
  double dist_bh1 = grav_dist(s, 1, r, tiph, 0.);
  double VFlux[16], AvgPrim[16], Grad_r_prim[16], Grad_ph_prim[16];
  for (int q = 0; q < NQ; ++q) {
    AvgPrim[q] = .5 * (R->primL[q] + R->primR[q]);
    Grad_ph_prim[q] = .5 * (GP(s, q, R->cL) + GP(s, q, R->cR)) / r;
    Grad_r_prim[q] = .5 * (G(s, q, R->cL) + G(s, q, R->cR));
    if (R->n[0] == 1) {
      double pL = s->tiph[R->cL] - .5 * s->dphi[R->cL];
      double pR = s->tiph[R->cR] - .5 * s->dphi[R->cR];
      double dpL = R->cm - pL;
      double dpR = -R->cm + pR;
      dpL = wrap_pm(dpL);
      dpR = wrap_pm(dpR);
      double WL = P(s, q, R->cL) + dpL * GP(s, q, R->cL);
      double WR = P(s, q, R->cR) - dpR * GP(s, q, R->cR);
      double deltaL = R->r - R->r_cell_L;
      double deltaR = R->r_cell_R - R->r;
      AvgPrim[q] = 0.5 * (WL + WR);
      Grad_r_prim[q] = (WR - WL) / (deltaR + deltaL);
    }
    if (R->n[1] == 1) {
      AvgPrim[q] = 0.5 * (P(s, q, R->cR) + P(s, q, R->cL));
      Grad_ph_prim[q] = (P(s, q, R->cR) - P(s, q, R->cL)) / (s->tiph[R->cR] - s->tiph[R->cL]) / r;
    }
    VFlux[q] = 0.0;
  }
  double nu;
  double alpha = s->dp[OD_EXPLICIT_VISCOSITY];
  double Gamma = s->dp[OD_GAMMALAW];
  if (s->ip[OI_VISC_CONST] == 1) {
    nu = alpha;
  } else if (s->ip[OI_InitialDataType] == SHEAR) {
    nu = alpha * Gamma * AvgPrim[PPP] / AvgPrim[RHO] * pow(fabs(r * cos(tiph)), 2.0);
    if (r * cos(tiph) > 20.) nu = 0.0;
  } else {
    double eps = s->dp[OD_G_EPS];
    nu = alpha * AvgPrim[PPP] / AvgPrim[RHO] /
         sqrt(pow(dist_bh0 * dist_bh0 + eps * eps, -1.5) * M0 + pow(dist_bh1 * dist_bh1 + eps * eps, -1.5) * M1);
  }
  double rho = AvgPrim[RHO];
  double vr = AvgPrim[URR];
  double om = AvgPrim[UPP];
  double vz = AvgPrim[UZZ];
  if (fabs(vz) > 1.e-12) return 12; /* the reference exits here (:271-274) */
  double Gr_vr = Grad_r_prim[URR];
  double Gr_om = Grad_r_prim[UPP];
  double Gp_vr = Grad_ph_prim[URR];
  double Gp_om = Grad_ph_prim[UPP];
  double rdr_om_cell = rdrOm_a(s, r, a);
  if (R->n[0] == 1) {
    VFlux[SRR] = -nu * rho * (r * Gr_vr - vr - r * r * Gp_om);
    VFlux[LLL] = -nu * rho * (r * (r * Gr_om + rdr_om_cell) + r * Gp_vr);
    VFlux[SZZ] = 0.0;
    VFlux[TAU] = -nu * rho * (vr * (Gr_vr - vr / r - r * Gp_om) + r * om * (Gp_vr + (rdr_om_cell + r * Gr_om)));
  }
  if (R->n[1] == 1) {
    VFlux[SRR] = -nu * rho * (r * (r * Gr_om + rdr_om_cell) + r * Gp_vr);
    VFlux[LLL] = -nu * rho * (-r * Gr_vr + vr + r * r * Gp_om);
    VFlux[SZZ] = 0.0;
    VFlux[TAU] = -nu * rho * (-r * om * (Gr_vr - vr / r - r * Gp_om) + vr * (Gp_vr + (rdr_om_cell + r * Gr_om)));
  }
  for (int q = 0; q < NQ; ++q) R->Fvisc[q] = VFlux[q];
  return 0;
}
static int add_flux(OracleSim *s, ORiemann *R, double dt) { /* Riemann_routines.c:314-397 */
  int NQ = s->NQ;
  double G = s->dp[OD_GAMMALAW];
  set_vel(R, R->r, G);
  double w = R->n[1] ? s->wiph[R->cL] : 0.0;
  set_state(R, w);
  if (R->state == LEFT) {
    set_flux(s, R, G, LEFT);
    prim2cons(s, R->primL, R->UL, R->r, 1.0);
    for (int q = 0; q < NQ; ++q) R->F[q] = R->FL[q] - w * R->UL[q];
  } else if (R->state == RIGHT) {
    set_flux(s, R, G, RIGHT);
    prim2cons(s, R->primR, R->UR, R->r, 1.0);
    for (int q = 0; q < NQ; ++q) R->F[q] = R->FR[q] - w * R->UR[q];
  } else {
    if (s->ip[OI_Riemann] == HLL) {
      set_flux(s, R, G, LEFT);
      set_flux(s, R, G, RIGHT);
      prim2cons(s, R->primL, R->UL, R->r, 1.0);
      prim2cons(s, R->primR, R->UR, R->r, 1.0);
      star_hll(s, R);
    } else if (s->ip[OI_Riemann] == HLLC) {
      if (R->state == LEFTSTAR) {
        set_flux(s, R, G, LEFT);
        prim2cons(s, R->primL, R->UL, R->r, 1.0);
      } else {
        set_flux(s, R, G, RIGHT);
        prim2cons(s, R->primR, R->UR, R->r, 1.0);
      }
      star_hllc(s, R, G);
    } else {
      return 13;
    }
    for (int q = 0; q < NQ; ++q) R->F[q] = R->Fstar[q] - w * R->Ustar[q];
  }
  if (s->dp[OD_EXPLICIT_VISCOSITY] > 0.0) {
    int e = visc_flux(s, R);
    if (e) return e;
  }
  for (int q = 0; q < NQ; ++q) {
    U(s, q, R->cL) += -dt * R->dA * R->F[q];
    U(s, q, R->cR) += dt * R->dA * R->F[q];
  }
  if (s->dp[OD_EXPLICIT_VISCOSITY] > 0.) {
    U(s, SRR, R->cL) += -dt * R->dA * R->Fvisc[SRR] / R->r_cell_L;
    U(s, SRR, R->cR) += dt * R->dA * R->Fvisc[SRR] / R->r_cell_R;
    U(s, LLL, R->cL) += -dt * R->dA * R->Fvisc[LLL];
    U(s, LLL, R->cR) += dt * R->dA * R->Fvisc[LLL];
    U(s, TAU, R->cL) += -dt * R->dA * R->Fvisc[TAU];
    U(s, TAU, R->cR) += +dt * R->dA * R->Fvisc[TAU];
  }
  return 0;
}

/* ---------------- sources: src/Cell/cell_source.c ---------------- */
static double fgrav(double M, double r, double eps, int n) { /* :11-13 */
  return (M * pow(r, n - 1.) / pow(pow(r, n) + pow(eps, n), 1. + 1. / n));
}
static void get_rho_sink(const OracleSim *s, int p, double dt, double r0, double r1, double M0, double M1,
                         double rho, double Pp, double *drho) { /* :15-47 */
  double Mtotal = M0 + M1;
  double alpha = s->dp[OD_EXPLICIT_VISCOSITY];
  double one_o_nu = 1.0 / (alpha * Pp / rho) * sqrt(pow(r0, -3) * M0 / Mtotal + pow(r1, -3) * M1 / Mtotal);
  double t_visc0 = TVISC_FAC * 2. / 3. * r0 * r0 * one_o_nu;
  double t_visc1 = TVISC_FAC * 2. / 3. * r1 * r1 * one_o_nu;
  if (t_visc0 < 10. * dt) t_visc0 = 10. * dt;
  if (t_visc1 < 10. * dt) t_visc1 = 10. * dt;
  *drho = 0.0;
  double sink_size = 0.5;
  if (p == 0) {
    if (r0 < sink_size) *drho = rho / t_visc0;
  } else if (p == 1) {
    if (r1 < sink_size) *drho = rho / t_visc1;
  }
}
static void grav_force(const OracleSim *s, int p, double r, double phi, double *fr, double *fp) { /* :49-71 */
  double G_EPS = s->dp[OD_G_EPS];
  int PHI_ORDER = s->ip[OI_PHI_ORDER];
  double rp = s->gm[p].r;
  double pp = s->gm[p].phi;
  double cosp = cos(phi);
  double sinp = sin(phi);
  double dx = r * cosp - rp * cos(pp);
  double dy = r * sinp - rp * sin(pp);
  double script_r = sqrt(dx * dx + dy * dy);
  double cosa = dx / script_r;
  double sina = dy / script_r;
  double cosap = cosa * cosp + sina * sinp;
  double sinap = sina * cosp - cosa * sinp;
  double f1 = -fgrav(s->gm[p].M, script_r, G_EPS, PHI_ORDER);
  *fr = cosap * f1;
  *fp = sinap * f1;
}
static void add_src(OracleSim *s, double dt) { /* :73-214 */
  int GRAV2D = s->ip[OI_GRAV2D];
  double Mdot_temp[2] = {0.0, 0.0};
  int imin_noghost = s->ip[OI_NgMinR] == 1 ? 0 : s->ip[OI_NgMinR];
  int kmin_noghost = s->ip[OI_NgMinZ];
  int imax_noghost = s->NR - s->ip[OI_NgMaxR];
  int kmax_noghost = s->NZ - s->ip[OI_NgMaxZ];
  double M0 = s->gm[0].M, M1 = s->gm[1].M;
  double rbh0 = s->gm[0].r, rbh1 = s->gm[1].r;
  double a = rbh0 + rbh1;
  double phi_bh0 = s->gm[0].phi, phi_bh1 = s->gm[1].phi;
  double xbh0 = rbh0 * cos(phi_bh0);
  double ybh0 = rbh0 * sin(phi_bh0);
  double xbh1 = rbh1 * cos(phi_bh1);
  double ybh1 = rbh1 * sin(phi_bh1);
  for (int k = 0; k < s->NZ; ++k) {
    double zm = zpos(s, k - 1), zp = zpos(s, k);
    for (int i = 0; i < s->NR; ++i) {
      double rm = rpos(s, i - 1), rp = rpos(s, i);
      for (int j = 0; j < s->N_p[i]; ++j) {
        int c = zid(s, k, i, j);
        double phi = s->tiph[c] - .5 * s->dphi[c];
        double dphi = s->dphi[c];
        double rho = P(s, RHO, c);
        double Pp = P(s, PPP, c);
        double r = .5 * (rp + rm);
        double vr = P(s, URR, c);
        double vz = P(s, UZZ, c);
        double vp = P(s, UPP, c) * r;
        double dz = zp - zm;
        double dV = dphi * .5 * (rp * rp - rm * rm) * dz;
        double xpos = r * cos(phi);
        double ypos = r * sin(phi);
        double r0 = sqrt((xpos - xbh0) * (xpos - xbh0) + (ypos - ybh0) * (ypos - ybh0));
        double r1 = sqrt((xpos - xbh1) * (xpos - xbh1) + (ypos - ybh1) * (ypos - ybh1));
        double z = .5 * (zp + zm);
        double R = sqrt(r * r + z * z);
        double sint, cost, gravdist;
        if (GRAV2D == 1) {
          sint = 1.0;
          cost = 0.0;
          gravdist = r;
        } else {
          sint = r / R;
          cost = z / R;
          gravdist = R;
        }
        double fr, fp, Fr = 0.0, Fp = 0.0;
        for (int p = 0; p < 2; ++p) {
          grav_force(s, p, gravdist, phi, &fr, &fp);
          Fr += fr;
          Fp += fp;
        }
        double w_a = rOm_a(s, r, a);
        double rdrOm = rdrOm_a(s, r, a);
        double F_centrifugal_r = w_a * w_a / r;
        double dtOm = dtOm_a(s, r, a);
        double F_euler_phi = -(r * dtOm + vr * rdrOm);
        U(s, SRR, c) += dt * dV * (rho * vp * vp + Pp) / r;
        U(s, SRR, c) += dt * dV * rho * (Fr * sint + F_centrifugal_r);
        U(s, LLL, c) += dt * dV * rho * (Fp - r * dtOm) * r;
        if (i >= s->ip[OI_NgMinR] && i < s->NR - s->ip[OI_NgMaxR] && k >= s->ip[OI_NgMinZ] &&
            k < s->NZ - s->ip[OI_NgMaxZ]) {
          s->torque += dt * dV * rho * Fp * r;
          s->led_torque += dt * dV * rho * Fp * r;
        }
        U(s, SZZ, c) += dt * dV * rho * Fr * cost;
        U(s, TAU, c) += dt * dV * rho * ((Fr * sint + F_centrifugal_r) * vr + Fr * vz * cost + (Fp + F_euler_phi) * vp);
        if (s->ip[OI_RhoSinkOn] == 1) {
          double drho;
          int num_sinks = s->ip[OI_GravMassType];
          for (int p = 0; p < num_sinks; ++p) {
            get_rho_sink(s, p, dt, r0, r1, M0, M1, rho, Pp, &drho);
            U(s, RHO, c) -= drho * dt * dV;
            if (i >= imin_noghost && i < imax_noghost && k >= kmin_noghost && k < kmax_noghost)
              Mdot_temp[p] += drho * dV;
          }
        }
        if (s->ip[OI_COOLING] == 1) {
          if (s->ip[OI_InitialDataType] == SHEAR) {
            double cooling_term = (Pp - .01) / (s->dp[OD_GAMMALAW] - 1.) * dV / 1.e-2;
            if (cooling_term < 0.0) cooling_term = 0.0;
            U(s, TAU, c) -= cooling_term * dt;
            s->Cool[c] = cooling_term / dV;
          } else {
            double numfac = 1.0;
            double alpha = s->dp[OD_EXPLICIT_VISCOSITY];
            double cooling_cap = U(s, TAU, c) * numfac;
            double cooling_term = 9. / 4. * alpha * pow(s->dp[OD_PoRho_r1], -3) / rho * pow(Pp / rho, 4.) * dt * dV;
            if (cooling_term > cooling_cap) cooling_term = cooling_cap;
            U(s, TAU, c) -= cooling_term;
            s->Cool[c] = cooling_term / dt / dV;
          }
        }
      }
    }
  }
  s->gm[0].Mdot = Mdot_temp[0];
  s->gm[1].Mdot = Mdot_temp[1];
}
static void add_visc_src(OracleSim *s, double dt) { /* :216-277 */
  double M0 = s->gm[0].M, M1 = s->gm[1].M;
  double r0 = s->gm[0].r, r1 = s->gm[1].r;
  double a = r0 + r1;
  double phi_bh0 = s->gm[0].phi, phi_bh1 = s->gm[1].phi;
  double xbh0 = r0 * cos(phi_bh0);
  double ybh0 = r0 * sin(phi_bh0);
  double xbh1 = r1 * cos(phi_bh1);
  double ybh1 = r1 * sin(phi_bh1);
  for (int k = 0; k < s->NZ; ++k) {
    double zm = zpos(s, k - 1), zp = zpos(s, k);
    for (int i = 0; i < s->NR; ++i) {
      double rm = rpos(s, i - 1), rp = rpos(s, i);
      for (int j = 0; j < s->N_p[i]; ++j) {
        int c = zid(s, k, i, j);
        double dphi = s->dphi[c];
        double rho = P(s, RHO, c);
        double Pp = P(s, PPP, c);
        double r = .5 * (rp + rm);
        double dz = zp - zm;
        double dV = dphi * .5 * (rp * rp - rm * rm) * dz;
        double phi = s->tiph[c] - 0.5 * s->dphi[c];
        double xpos = r * cos(phi);
        double ypos = r * sin(phi);
        double dist_bh0 = sqrt((xpos - xbh0) * (xpos - xbh0) + (ypos - ybh0) * (ypos - ybh0));
        double dist_bh1 = sqrt((xpos - xbh1) * (xpos - xbh1) + (ypos - ybh1) * (ypos - ybh1));
        double alpha = s->dp[OD_EXPLICIT_VISCOSITY];
        double rdrOm = rdrOm_a(s, r, a);
        double eps = s->dp[OD_G_EPS];
        double Sigma_nu;
        if (s->ip[OI_VISC_CONST] == 1)
          Sigma_nu = rho * alpha;
        else
          Sigma_nu = alpha * Pp / sqrt(pow(dist_bh0 * dist_bh0 + eps * eps, -1.5) * M0 +
                                       pow(dist_bh1 * dist_bh1 + eps * eps, -1.5) * M1);
        U(s, TAU, c) += dt * dV * Sigma_nu * rdrOm * (rdrOm + r * G(s, UPP, c) + GP(s, URR, c));
      }
    }
  }
}
static void split_fictitious(OracleSim *s, double dt) { /* cell_add_split_fictitious.c:11-44 */
  double a = s->gm[0].r + s->gm[1].r;
  for (int k = 0; k < s->NZ; ++k)
    for (int i = 0; i < s->NR; ++i) {
      double rm = rpos(s, i - 1), rp = rpos(s, i);
      double r = 0.5 * (rm + rp);
      for (int j = 0; j < s->N_p[i]; ++j) {
        int c = zid(s, k, i, j);
        double Omega = rOm_a(s, r, a) / r;
        double dr_r2Om = 2. * rOm_a(s, r, a) + r * rdrOm_a(s, r, a);
        double kappa2 = 2. * Omega / r * dr_r2Om;
        if (kappa2 > 0.) {
          double kappa = sqrt(kappa2);
          double l0 = U(s, LLL, c);
          double rhov0 = U(s, SRR, c);
          double l = l0 * cos(kappa * dt) - U(s, SRR, c) * dr_r2Om / kappa * sin(kappa * dt);
          double rhov = rhov0 * cos(kappa * dt) + 2.0 * Omega * U(s, LLL, c) / r / kappa * sin(kappa * dt);
          U(s, LLL, c) = l;
          U(s, SRR, c) = rhov;
        }
      }
    }
}

/* ---------------- boundaries: src/Cell/cell_boundary.c:14-139 ---------------- */
static void outflow_r(OracleSim *s) {
  int NR = s->NR, NQ = s->NQ;
  OFace *f = s->faces[0];
  int Nf = s->nri[NR - 1];
  int n1 = s->nri[NR - 2];
  int n2 = s->nri[NR - 3];
  if (s->ip[OI_NoInnerBC] != 1) {
    int i = 0;
    double r_face = rpos(s, i);
    for (int n = s->nri[i]; n < s->nri[i + 1]; ++n)
      for (int q = 0; q < NQ; ++q) P(s, q, f[n].L) = 0.0;
    for (int n = s->nri[i]; n < s->nri[i + 1]; ++n) {
      int cL = f[n].L, cR = f[n].R;
      for (int q = 0; q < NQ; ++q)
        P(s, q, cL) += (P(s, q, cR) - 0 * G(s, q, cR) * (rpos(s, i + 1) - rpos(s, i - 1)) / 2.) * f[n].dA;
    }
    for (int k = 0; k < s->NZ; ++k) {
      double zp = zpos(s, k), zm = zpos(s, k - 1);
      double dz = zp - zm;
      for (int j = 0; j < s->N_p[i]; ++j) {
        int c = zid(s, k, i, j);
        double dA = dz * r_face * s->dphi[c];
        for (int q = 0; q < NQ; ++q) P(s, q, c) /= dA;
      }
    }
  }
  for (int n = n2; n < n1; ++n)
    for (int q = 0; q < NQ; ++q) P(s, q, f[n].R) = 0.0;
  for (int n = n2; n < n1; ++n)
    for (int q = 0; q < NQ; ++q) P(s, q, f[n].R) += P(s, q, f[n].L) * f[n].dA;
  for (int k = 0; k < s->NZ; ++k) {
    double zp = zpos(s, k), zm = zpos(s, k - 1);
    double dz = zp - zm;
    for (int j = 0; j < s->N_p[NR - 2]; ++j) {
      double r_face = rpos(s, NR - 3);
      int c = zid(s, k, NR - 2, j);
      double dA = dz * r_face * s->dphi[c];
      for (int q = 0; q < NQ; ++q) P(s, q, c) /= dA;
    }
  }
  for (int n = n1; n < Nf; ++n)
    for (int q = 0; q < NQ; ++q) P(s, q, f[n].R) = 0.0;
  for (int n = n1; n < Nf; ++n)
    for (int q = 0; q < NQ; ++q) P(s, q, f[n].R) += P(s, q, f[n].L) * f[n].dA;
  for (int k = 0; k < s->NZ; ++k) {
    double zp = zpos(s, k), zm = zpos(s, k - 1);
    double dz = zp - zm;
    for (int j = 0; j < s->N_p[NR - 1]; ++j) {
      double r_face = rpos(s, NR - 2);
      int c = zid(s, k, NR - 1, j);
      double dA = dz * r_face * s->dphi[c];
      for (int q = 0; q < NQ; ++q) P(s, q, c) /= dA;
    }
  }
}

/* z halo on one rank: periodic copy of (prim, RKcons, tiph), src/Cell/cell_sync.c:156-205 */
static void sync_z(OracleSim *s) {
  if (s->ip[OI_NglobalZ] <= 1) return;
  int NZ = s->NZ, ngmax = s->ip[OI_NgMaxZ], ngmin = s->ip[OI_NgMinZ];
  int kNm2g = NZ - 2 * ngmax, kNmg = NZ - ngmax, kg = ngmin;
  for (int d = 0; d < ngmax; ++d) { /* upper ghosts <- lower interior */
    int kdst = kNmg + d, ksrc = kg + d;
    for (int x = 0; x < s->NL; ++x) {
      int zd = kdst * s->NL + x, zs = ksrc * s->NL + x;
      for (int q = 0; q < s->NQ; ++q) {
        P(s, q, zd) = P(s, q, zs);
        RU(s, q, zd) = RU(s, q, zs);
      }
      s->tiph[zd] = s->tiph[zs];
    }
  }
  for (int d = 0; d < ngmin; ++d) { /* lower ghosts <- upper interior */
    int kdst = d, ksrc = kNm2g + d;
    for (int x = 0; x < s->NL; ++x) {
      int zd = kdst * s->NL + x, zs = ksrc * s->NL + x;
      for (int q = 0; q < s->NQ; ++q) {
        P(s, q, zd) = P(s, q, zs);
        RU(s, q, zd) = RU(s, q, zs);
      }
      s->tiph[zd] = s->tiph[zs];
    }
  }
}

/* ---------------- initial data: src/Cell/initial_data/cell_init_*.c -------
 * The seven plugins of src/Cell/cell_init_function_ptr.c:6-46.  The global
 * functions write the cells' primitives; the single-cell ones write one
 * cell, given its annulus i and edges.  Both keep the reference's loops and
 * expression order; the torus draws from the C library's own rand(). */
enum { ID_SHEAR, ID_VORTEX, ID_SSTEST, ID_TORUS, ID_MILOS, ID_RAD_DOM, ID_MIDDLE };

static double ann_r(const OracleSim *s, int i) { return 0.5 * (rpos(s, i - 1) + rpos(s, i)); }

/* the five primitives of the single-cell function of the disk's type for a
 * cell of annulus i with edge t_iph and width d_phi (cell_single_init_*):
 * returns 1 when it writes them, 0 when it writes nothing (the vortex
 * outside its core), -1 for the torus (cell_init_torus.c:12-15 exits);
 * *passive gets the vortex's passive value when there is one */
static int single_init(const OracleSim *s, int i, double t_iph, double d_phi, double *pr, double *passive) {
  double r = ann_r(s, i);
  switch (s->ip[OI_InitialDataType]) {
    case ID_SHEAR: { /* cell_init_shear.c:11-55 */
      double rho = 1.0, Pp = 0.01, v0 = 1.0, t0 = 0.25;
      double t = t_iph - .5 * d_phi;
      double x = r * cos(t) - 3.;
      double y = r * sin(t);
      double nu = s->dp[OD_EXPLICIT_VISCOSITY] > 0.0 ? s->dp[OD_EXPLICIT_VISCOSITY] : 0.05;
      double vy = 0.0;
      if (fabs(y) < 50.) vy = v0 / sqrt(2. * M_PI * nu * t0) * exp(-x * x / (4. * nu * t0));
      double vr = vy * sin(t);
      double omega = vy * cos(t) / r;
      pr[RHO] = rho; pr[PPP] = Pp; pr[URR] = vr; pr[UPP] = omega - rOm_a(s, r, 1.) / r; pr[UZZ] = 0.0;
      return 1;
    }
    case ID_VORTEX: { /* cell_init_vortex.c:128-154 */
      double rho = 1.0, P0 = 1.0, R = 1.0, A = 0.5, g = 3.0;
      double t = t_iph - .5 * d_phi;
      if (r >= R) return 0;
      double Pp = P0 * (1. - A * exp(g * (r / R - R / (R - r) + 1.)));
      double vp = sqrt(A * g * P0 / rho * (2. * R - r) / R * r * r / pow(R - r, 2.) * exp(g * (r / R - R / (R - r) + 1.)));
      *passive = cos(t) < 0.0 ? 0.0 : 1.0;
      pr[RHO] = rho; pr[PPP] = Pp; pr[URR] = 0.0; pr[UPP] = vp / r; pr[UZZ] = 0.0;
      return 1;
    }
    case ID_SSTEST:
    case ID_MIDDLE: { /* cell_init_SStest.c:12-47, cell_init_middle.c:12-47 */
      int mid = s->ip[OI_InitialDataType] == ID_MIDDLE;
      double alpha = s->dp[OD_EXPLICIT_VISCOSITY];
      double r0 = mid ? 2.5 : 5.;
      int d = mid ? 10 : 4;
      double rho = pow(r, -3. / 5.) * exp(-pow(r / r0, -d));
      double P = s->dp[OD_PoRho_r1] * pow(r, -3. / 2.) * exp(-pow(r / r0, -d));
      double r_drP_o_P = -1.5 + d * pow(r / r0, -d);
      if (rho < (mid ? 1.e-3 : 1.e-4)) rho = mid ? 1.e-3 : 1.e-4;
      if (P < (mid ? 1.e-6 : 1.e-7)) P = mid ? 1.e-6 : 1.e-7;
      double o2 = 1. / r / r / r * pow(1. + 3. / 16. / r / r, 2) + r_drP_o_P * P / rho / r / r;
      double omega = sqrt(o2);
      double a = 1.0;
      omega = (pow(r, W_A_MILOS_INDEX) * omega + pow(a, W_A_MILOS_INDEX - 1.5)) /
              (pow(r, W_A_MILOS_INDEX) + pow(a, W_A_MILOS_INDEX));
      double vr = mid ? -3. * alpha / r / omega * P / rho * (2. + r_drP_o_P) : 0.0;
      pr[RHO] = rho; pr[PPP] = P; pr[URR] = vr;
      pr[UPP] = (mid ? omega : pow(r, -1.5)) - rOm_a(s, r, a) / r;
      pr[UZZ] = 0.0;
      return 1;
    }
    case ID_MILOS: { /* cell_init_milos_macfadyen.c:11-35 */
      double Mach = 1. / sqrt(s->dp[OD_PoRho_r1]);
      double Gam = s->dp[OD_GAMMALAW];
      double rho0 = 1.0, rs = 10., delta_exp = 3.0, xi_exp = 2.0;
      double rho = rho0 * (pow(rs / r, delta_exp)) * exp(-pow(rs / r, xi_exp));
      double PoRho = 1. / Mach / Mach / Gam;
      pr[RHO] = rho; pr[PPP] = PoRho * rho; pr[URR] = 0.0; pr[UPP] = 0.0; pr[UZZ] = 0.0;
      return 1;
    }
    case ID_RAD_DOM: { /* cell_init_rad_dom.c:103-131 */
      double blah = 0.0;
      double xi = 0.005, a_o_M = 100;
      double alpha = s->dp[OD_EXPLICIT_VISCOSITY];
      double Gam = s->dp[OD_GAMMALAW];
      double rho0 = 2. / (9. * xi * alpha * Gam) * sqrt(a_o_M);
      double P0 = 8. * xi / (9. * alpha * Gam) * sqrt(a_o_M);
      double rho = rho0 * pow(r, 1.5);
      double P = P0 * pow(r, -1.5);
      double o2 = 1. / r / r / r - 3. / 2. * P / rho / r / r;
      double omega = sqrt(o2);
      double vr = -1.5 * alpha * Gam * (P / rho) * sqrt(r) - blah * pow(r, -4.5);
      pr[RHO] = rho; pr[PPP] = P; pr[URR] = vr; pr[UPP] = omega - 1. / pow(r, 1.5); pr[UZZ] = 0.0;
      return 1;
    }
  }
  return -1;
}

/* the single-cell function applied to cell z of annulus i; it also writes
 * the cell's wiph, which cell_clear_w resets before any use */
static int single_cell(OracleSim *s, int i, int z) {
  double pr[5], passive = 0.0;
  int type = s->ip[OI_InitialDataType];
  int set = single_init(s, i, s->tiph[z], s->dphi[z], pr, &passive);
  if (set < 0) return 17;
  if (set == 0) return 0;
  for (int q = 0; q < 5; ++q) P(s, q, z) = pr[q];
  if (type == ID_VORTEX && s->NC < s->NQ) P(s, s->NC, z) = passive;
  if (type == ID_RAD_DOM)
    s->wiph[z] = pow(ann_r(s, i), -.5); /* cell_init_rad_dom.c:128 */
  else if (type != ID_VORTEX)
    s->wiph[z] = 0.0;
  return 0;
}

static int fixed_r(OracleSim *s) { /* cell_boundary_fixed_r, cell_boundary.c:208-232 */
  int e;
  if (s->ip[OI_NoInnerBC] != 1)
    for (int i = 0; i < s->ip[OI_NgMinR]; ++i)
      for (int k = 0; k < s->NZ; ++k)
        for (int j = 0; j < s->N_p[i]; ++j)
          if ((e = single_cell(s, i, zid(s, k, i, j)))) return e;
  for (int i = s->NR - 1; i > s->NR - s->ip[OI_NgMaxR] - 1; --i)
    for (int k = 0; k < s->NZ; ++k)
      for (int j = 0; j < s->N_p[i]; ++j)
        if ((e = single_cell(s, i, zid(s, k, i, j)))) return e;
  return 0;
}

static int fixed_z(OracleSim *s) { /* cell_boundary_fixed_z, cell_boundary.c:234-254 */
  int e;
  for (int i = 0; i < s->NR; ++i)
    for (int k = 0; k < s->ip[OI_NgMinZ]; ++k)
      for (int j = 0; j < s->N_p[i]; ++j)
        if ((e = single_cell(s, i, zid(s, k, i, j)))) return e;
  for (int i = 0; i < s->NR; ++i)
    for (int k = s->NZ - 1; k > s->NZ - s->ip[OI_NgMaxZ] - 1; --k)
      for (int j = 0; j < s->N_p[i]; ++j)
        if ((e = single_cell(s, i, zid(s, k, i, j)))) return e;
  return 0;
}

/* cell_bc_damp, cell_misc.c:98-132.  The reference's initialCell comes from
 * an uninitialised malloc (cell_create_destroy.c:12-16): what the plugin does
 * not write, and the cell's edges, are NaN here. */
static int bc_damp(OracleSim *s, double dt) {
  double RMIN = s->dp[OD_RMIN], RMAX = s->dp[OD_RMAX];
  double RDI = s->dp[OD_RDAMP_INNER], RDO = s->dp[OD_RDAMP_OUTER];
  double R0 = 1. / s->dp[OD_DAMP_TIME] / 2. / M_PI;
  double init[16];
  for (int i = 0; i < s->NR; ++i) {
    double rp = rpos(s, i), rm = rpos(s, i - 1);
    double r = .5 * (rp + rm);
    if (r < RDI || r > RDO) {
      double rate = R0 * (r - RDI) / (RMIN - RDI);
      if (r > RDO) rate = R0 * (r - RDO) / (RMAX - RDO);
      for (int k = 0; k < s->NZ; ++k)
        for (int j = 0; j < s->N_p[i]; ++j) {
          double passive = NAN;
          for (int q = 0; q < s->NQ; ++q) init[q] = NAN;
          if (single_init(s, i, NAN, NAN, init, &passive) < 0) return 17;
          if (s->ip[OI_InitialDataType] == ID_VORTEX && s->NC < s->NQ) init[s->NC] = passive;
          int z = zid(s, k, i, j);
          for (int q = 0; q < s->NQ; ++q) {
            double dprim = init[q] - P(s, q, z);
            P(s, q, z) += (1. - exp(-rate * dt)) * dprim;
          }
        }
    }
  }
  return 0;
}

/* the global functions, cell_init_*.c; returns 0, or 17 for an unknown type */
int oracle_initial_data(OracleSim *s) {
  grav_init(s); /* src/main.c:103 sets the masses first */
  int type = s->ip[OI_InitialDataType];
  if (type == ID_TORUS) { /* cell_init_torus.c:17-90 */
    double GAMMALAW = s->dp[OD_GAMMALAW];
    double M = 1.0, r_inner = 0.7, r_max = 1.0, q = 2.;
    double sx = 2. * q - q;
    double GammaFac = (GAMMALAW - 1.) / GAMMALAW;
    double Phi_inner = -1.0 / r_inner;
    double Phi_max = -1.0 / r_max;
    double Rho_norm = pow(GammaFac * (Phi_inner - Phi_max + pow(r_inner, -sx) / sx - pow(1.0, -sx) / sx),
                          1.0 / (GAMMALAW - 1.0));
    srand(666);
    for (int k = 0; k < s->NZ; ++k) {
      double z = s->ip[OI_NglobalZ] != 1 ? 0.5 * (zpos(s, k - 1) + zpos(s, k)) : 0.0;
      for (int i = 0; i < s->NR; ++i) {
        double r = ann_r(s, i);
        double R = sqrt(r * r + z * z);
        double Phi = -M / R;
        double omega = pow(r, -q);
        double PoRho = GammaFac * (Phi_inner - Phi + pow(r_inner, -sx) / sx - pow(r, -sx) / sx);
        if (PoRho < 1.0e-5) PoRho = 1.0e-5;
        double rho = pow(PoRho, 1.0 / (GAMMALAW - 1.0)) / Rho_norm;
        double Pp = pow(PoRho, 1.0 / GammaFac) / Rho_norm;
        if (PoRho < 1.0e-5) Pp = rho;
        for (int j = 0; j < s->N_p[i]; ++j) {
          double delta = .02 * ((double)rand() / (double)RAND_MAX - .5);
          int c = zid(s, k, i, j);
          P(s, RHO, c) = rho; P(s, PPP, c) = Pp; P(s, URR, c) = 0.0;
          P(s, UPP, c) = omega * (1. + delta) - pow(r, -2.);
          P(s, UZZ, c) = 0.0;
        }
      }
    }
    return 0;
  }
  if (type == ID_MILOS) { /* cell_init_milos_macfadyen.c:37-152 */
    double rs = 10.0, rho0 = 1.0;
    double Mach = 1. / sqrt(s->dp[OD_PoRho_r1]);
    double delta_exp = 3.0, xi_exp = 2.0;
    double Gam = s->dp[OD_GAMMALAW];
    double cp = 1. / Mach, fac = 0.0001, Mtotal = 1.0, sep = 1.0;
    double massratio = s->dp[OD_MassRatio];
    double M0, M1, r0, r1;
    if (fabs(massratio) > 1.e-8) {
      M0 = Mtotal / (1. + massratio);
      M1 = Mtotal / (1. + 1. / massratio);
      r0 = M1 / Mtotal * sep;
      r1 = M0 / Mtotal * sep;
    } else {
      M0 = 1.0; M1 = 0.0; r0 = 0.0; r1 = 0.0;
    }
    double eps0 = s->dp[OD_G_EPS] * r0, eps1 = s->dp[OD_G_EPS] * r1;
    double DISK_ALPHA = s->dp[OD_EXPLICIT_VISCOSITY];
    for (int k = 0; k < s->NZ; ++k)
      for (int i = 0; i < s->NR; ++i) {
        double r = ann_r(s, i);
        double cs = r < 1. ? 1. / Mach : sqrt(1. / r) / Mach;
        for (int j = 0; j < s->N_p[i]; ++j) {
          int c = zid(s, k, i, j);
          double phi = s->tiph[c] - 0.5 * s->dphi[c];
          double d0 = grav_dist(s, 0, r, phi, 0.), d1 = grav_dist(s, 1, r, phi, 0.);
          double cs0 = d0 < 0.5 ? sqrt(1. / .5) / Mach * sqrt(M0) : sqrt(1. / d0) / Mach * sqrt(M0);
          double cs1 = d1 < 0.5 ? sqrt(1. / .5) / Mach * sqrt(M1) : sqrt(1. / d1) / Mach * sqrt(M1);
          double PoRho = (cs0 * cs0 + cs1 * cs1) / Gam;
          int n = s->ip[OI_PHI_ORDER];
          double Pot1 = M0 / pow(pow(d0, n) + pow(eps0, n), 1. / n);
          double Pot2 = M1 / pow(pow(d1, n) + pow(eps1, n), 1. / n);
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
            vr = -3.0 / sqrt(r) * DISK_ALPHA * (1.0 / Mach) * (1.0 / Mach) * (1.0 - delta_exp + xi_exp * pow(r / rs, -xi_exp));
          }
          if (rho < (100. * s->dp[OD_RHO_FLOOR])) rho = 100. * s->dp[OD_RHO_FLOOR];
          double Pp = PoRho * rho;
          P(s, RHO, c) = rho * exp(fac * (Pot1 + Pot2) / cp / cp);
          P(s, PPP, c) = Pp * exp(fac * (Pot1 + Pot2) / cp / cp);
          P(s, URR, c) = vr;
          P(s, UPP, c) = omega - sqrt(1.) / pow(r, 1.5);
          P(s, UZZ, c) = 0.0;
        }
      }
    return 0;
  }
  if (type < ID_SHEAR || type > ID_MIDDLE) return 17;
  /* shear, vortex, rad_dom: the global loops evaluate what the single-cell
   * functions do; SStest and middle also zero v_r for r < 1 (:81) */
  for (int k = 0; k < s->NZ; ++k)
    for (int i = 0; i < s->NR; ++i)
      for (int j = 0; j < s->N_p[i]; ++j) {
        int c = zid(s, k, i, j);
        if (single_cell(s, i, c)) return 17;
        if ((type == ID_SSTEST || type == ID_MIDDLE) && ann_r(s, i) < 1.) P(s, URR, c) = 0.0;
      }
  return 0;
}

/* ---------------- driver ---------------- */
void oracle_init(OracleSim *s) { /* src/main.c:103-119 (IC plugin replaced by the caller's prim) */
  grav_init(s);
  grav_clean_pi(s);
  sync_z(s);
  calc_cons(s);
}

static int substep(OracleSim *s, double fac) { /* src/TimeStep/timestep_routines.c:16-113 */
  double dt = fac * s->dt;
  int e;
  oracle_face_create(s, 0);
  oracle_face_create(s, 1);
  clean_pi(s);
  grav_clean_pi(s);
  oracle_set_w(s);
  if (s->err) return s->err;
  adjust_RK_cons(s, s->RK);
  oracle_plm_p(s);
  oracle_plm_rz(s, 0);
  ORiemann R;
  for (int k = 0; k < s->NZ; ++k)
    for (int i = 0; i < s->NR; ++i)
      for (int j = 0; j < s->N_p[i]; ++j) {
        riemann_init(&R);
        setup_p(s, &R, i, j, k);
        if ((e = add_flux(s, &R, dt))) return e;
      }
  memset(s->led, 0, sizeof(s->led));
  s->led_torque = 0.0;
  const int in0 = s->nri[0], in1 = s->nri[1];
  const int out0 = s->nri[s->NR - 3], out1 = s->nri[s->NR - 2];
  for (int n = 0; n < s->nri[s->NR - 1]; ++n) {
    riemann_init(&R);
    setup_rz(s, &R, 0, n, 0);
    if ((e = add_flux(s, &R, dt))) return e;
    int side = (n >= in0 && n < in1) ? 0 : (n >= out0 && n < out1) ? 1 : -1;
    if (side >= 0) { /* test infrastructure: the conservation ledger */
      for (int q = 0; q < s->NQ; ++q) s->led[side][q] += dt * R.dA * R.F[q];
      if (s->dp[OD_EXPLICIT_VISCOSITY] > 0.)
        for (int v = 0; v < 3; ++v) s->led[side][8 + v] += dt * R.dA * R.Fvisc[v == 0 ? SRR : v == 1 ? LLL : TAU];
    }
  }
  if (s->ip[OI_NglobalZ] != 1) {
    oracle_plm_rz(s, 1);
    for (int n = 0; n < s->nzk[s->NZ - 1]; ++n) {
      riemann_init(&R);
      setup_rz(s, &R, 1, n, 2);
      if ((e = add_flux(s, &R, dt))) return e;
    }
  }
  add_src(s, dt);
  if (s->dp[OD_EXPLICIT_VISCOSITY] > 0.0) add_visc_src(s, dt);
  split_fictitious(s, dt);
  update_phi(s, s->RK, dt);
  update_dphi(s);
  grav_update_RK(s, s->RK);
  calc_prim(s);
  if (s->ip[OI_BoundTypeR] == BOUND_OUTFLOW)
    outflow_r(s);
  else if (s->ip[OI_BoundTypeR] == BOUND_FIXED && (e = fixed_r(s)))
    return e;
  /* outflow_z (cell_boundary.c:141-206) and fixed_z only write the ghost
   * layers, which the single-rank periodic z sync below overwrites. */
  if (s->ip[OI_NglobalZ] > 1 && s->ip[OI_BoundTypeZ] == BOUND_FIXED && (e = fixed_z(s))) return e;
  if (s->dp[OD_DAMP_TIME] > 0.0 && (e = bc_damp(s, dt))) return e;
  sync_z(s);
  calc_cons(s);
  return 0;
}

int oracle_rk2(OracleSim *s) { /* src/TimeStep/timestep_algorithm.c:11-41 */
  int e;
  oracle_set_w(s);
  if (s->err) return s->err;
  s->dt = oracle_mindt(s); /* timestep_set_dt, timestep_set.c:16-21 */
  if (s->t + s->dt > s->dp[OD_T_MAX]) s->dt = s->dp[OD_T_MAX] - s->t;
  cell_copy(s);
  grav_copy(s);
  s->RK = 0.0;
  if ((e = substep(s, 1.0))) return s->err = e;
  grav_move(s, s->t, 1.0 * s->dt);
  s->RK = 0.5;
  if ((e = substep(s, 0.5))) return s->err = e;
  grav_move(s, s->t, 0.5 * s->dt);
  s->t += s->dt;
  for (int p = 0; p < 2; ++p) s->gm[p].Macc = s->gm[p].Macc + s->gm[p].Mdot * s->dt;
  return 0;
}

int oracle_forward_euler(OracleSim *s) { /* src/TimeStep/timestep_algorithm.c:43-57 */
  int e;
  oracle_set_w(s);
  if (s->err) return s->err;
  s->dt = oracle_mindt(s); /* timestep_set_dt */
  if (s->t + s->dt > s->dp[OD_T_MAX]) s->dt = s->dp[OD_T_MAX] - s->t;
  cell_copy(s);
  grav_copy(s);
  s->RK = 0.0;
  if ((e = substep(s, 1.0))) return s->err = e;
  grav_move(s, s->t, 1.0 * s->dt);
  s->t += s->dt; /* timestep_update_t; no Macc update */
  return 0;
}

/* timestep_rk2 in pieces (test infrastructure for the conservation
 * ledger): the step's set-up, one substep with its mass move, the end */
int oracle_step_begin(OracleSim *s) {
  oracle_set_w(s);
  if (s->err) return s->err;
  s->dt = oracle_mindt(s);
  if (s->t + s->dt > s->dp[OD_T_MAX]) s->dt = s->dp[OD_T_MAX] - s->t;
  cell_copy(s);
  grav_copy(s);
  s->RK = 0.0;
  return 0;
}
int oracle_step_substep(OracleSim *s, double fac) {
  int e = substep(s, fac);
  if (e) return s->err = e;
  grav_move(s, s->t, fac * s->dt);
  s->RK = 0.5;
  return 0;
}
void oracle_step_end(OracleSim *s) {
  s->t += s->dt;
  for (int p = 0; p < 2; ++p) s->gm[p].Macc = s->gm[p].Macc + s->gm[p].Mdot * s->dt;
}
/* the last substep's ledger: led[0], led[1] (NQ + 3 each), torque */
void oracle_ledger(OracleSim *s, double *out) {
  const int n = s->NQ + 3;
  for (int v = 0; v < n; ++v) {
    out[v] = s->led[0][v < s->NQ ? v : 8 + (v - s->NQ)];
    out[n + v] = s->led[1][v < s->NQ ? v : 8 + (v - s->NQ)];
  }
  out[2 * n] = s->led_torque;
}

int oracle_run(OracleSim *s, int nsteps) {
  for (int n = 0; n < nsteps; ++n) {
    int e = oracle_rk2(s);
    if (e) return e;
  }
  return 0;
}

/* accumulated torque term (test infrastructure for the adjoint's loss) */
double oracle_torque(OracleSim *s) { return s->torque; }

/* ---------------- diagnostics: src/Diagnostics/diagnostics_set.c ---------------- */
/* number of equatorial cells, diagnostics_create_destroy.c:34-51 */
int oracle_diag_neq(OracleSim *s) {
  int n = 0;
  const int imin = s->ip[OI_NgMinR], imax = s->NR - s->ip[OI_NgMaxR];
  const int kmin = s->ip[OI_NgMinZ], kmax = s->NZ - s->ip[OI_NgMaxZ];
  for (int k = kmin; k < kmax; ++k) {
    double zp = zpos(s, k), zm = zpos(s, k - 1), z = 0.5 * (zm + zp);
    for (int i = imin; i < imax; ++i)
      for (int j = 0; j < s->N_p[i]; ++j)
        if ((fabs(zp) < 0.0000001) || (fabs(z) < 0.0000001)) ++n;
  }
  return n;
}

/* One measurement of diagnostics_set (diagnostics_set.c:13-303) on the
 * current state, one rank.  dtout = t - toutprev; rz = {R_Min, R_Max, Z_Min,
 * Z_Max} (sim_MIN/sim_MAX).  Outputs, as the reference forms them before
 * adding them to its running sums:
 *   vec[NRg * 17]   VectorDiag_reduce after its dtout/(ZMAX-ZMIN) scaling
 *   scal[16]        ScalarDiag_reduce after its scaling
 *   equat[Neq * 18] EquatDiag
 *   mdot[12]        the DiagMdot.dat line after its t: Mdot_near_req1,
 *                   r_near_req1, mass_inside_1, mass_inner_wedge, Mdot0,
 *                   Mdot1, Macc0, Macc1, masses within 0.2 and 0.4 of each mass
 * The passive scalar column reads prim[5]; with NUM_Q = 5 the reference reads
 * past the primitive array there, this restatement writes 0. */
void oracle_diag_measure(OracleSim *s, double dtout, const double *rz, double *vec, double *scal,
                         double *equat, double *mdot) {
  const int NV = 17, NE = 18, NS = 16;
  const int imin = s->ip[OI_NgMinR], imax = s->NR - s->ip[OI_NgMaxR];
  const int kmin = s->ip[OI_NgMinZ], kmax = s->NZ - s->ip[OI_NgMaxZ];
  const int nrg = imax - imin;
  const double G = s->dp[OD_GAMMALAW];
  double m01[2] = {0, 0}, m02[2] = {0, 0}, m04[2] = {0, 0}, inside1 = 0.0;
  for (int n = 0; n < nrg * NV; ++n) vec[n] = 0.0;
  for (int n = 0; n < NS; ++n) scal[n] = 0.0;
  int position = 0;
  for (int k = kmin; k < kmax; ++k) {
    double zp = zpos(s, k), zm = zpos(s, k - 1);
    double z = 0.5 * (zm + zp), dz = zp - zm;
    for (int i = imin; i < imax; ++i) {
      double rp = rpos(s, i), rm = rpos(s, i - 1);
      double r = 0.5 * (rm + rp);
      double a = s->gm[0].r + s->gm[1].r;
      for (int j = 0; j < s->N_p[i]; ++j) {
        int c = zid(s, k, i, j);
        double dphi = s->dphi[c];
        double phi = s->tiph[c] - 0.5 * dphi;
        double rho = s->prim[RHO * s->Nz + c];
        double press = s->prim[PPP * s->Nz + c];
        double vr = s->prim[URR * s->Nz + c];
        double vp_minus_w = s->prim[UPP * s->Nz + c] * r;
        double vp = s->prim[UPP * s->Nz + c] * r + rOm_a(s, r, a);
        double vz = s->prim[UZZ * s->Nz + c];
        double v2 = vr * vr + vp * vp + vz * vz;
        double rhoe = press / (G - 1.);
        double KE = 0.5 * rho * v2;
        double E_hydro = rhoe + KE;
        double Cool = s->Cool[c];
        double dV = 0.5 * (rp * rp - rm * rm) * dphi;
        if (rp <= 1.0) inside1 += rho * dV;
        double passive = s->NQ > 5 ? s->prim[5 * s->Nz + c] : 0.0;
        double d0 = grav_dist(s, 0, r, phi, 0.), d1 = grav_dist(s, 1, r, phi, 0.);
        if (d0 < 0.1) m01[0] += rho * dV;
        if (d1 < 0.1) m01[1] += rho * dV;
        if (d0 < 0.2) m02[0] += rho * dV;
        if (d1 < 0.2) m02[1] += rho * dV;
        if (d0 < 0.4) m04[0] += rho * dV;
        if (d1 < 0.4) m04[1] += rho * dV;
        if ((fabs(zp) < 0.0000001) || (fabs(z) < 0.0000001)) {
          double *e = equat + (size_t)position * NE;
          e[0] = r; e[1] = phi; e[2] = rho; e[3] = press; e[4] = vr; e[5] = vp; e[6] = vz;
          e[7] = vp_minus_w; e[8] = Cool; e[9] = passive; e[10] = E_hydro; e[11] = KE; e[12] = rhoe;
          e[13] = e[14] = e[15] = e[16] = e[17] = 0.0;
          ++position;
        }
        double *v = vec + (size_t)(i - imin) * NV;
        double np = s->N_p[i];
        v[0] += (r / np * dz);
        v[1] += (rho / np * dz);
        v[2] += (press / np * dz);
        v[3] += (vr / np * dz);
        v[4] += (vp / np * dz);
        v[5] += (vz / np * dz);
        v[6] += (vp_minus_w / np * dz);
        v[7] += (Cool / np * dz);
        v[8] += (passive / np * dz);
        v[9] += (E_hydro / np * dz);
        v[10] += (KE / np * dz);
        v[11] += (rhoe / np * dz);
        for (int n = 12; n < NV; ++n) v[n] += 0.0;
      }
    }
  }
  int i;
  for (i = imin; i < imax; ++i) {
    double rp = rpos(s, i), rm = rpos(s, i - 1);
    for (int n = 0; n < NS; ++n) scal[n] += vec[(size_t)(i - imin) * NV + n + 1] * (rp * rp - rm * rm);
  }
  /* :201-211: the loop bound is N_p of the index the previous loop left (imax),
   * and cell (0, j, 0) runs on through the layer's contiguous cells */
  double wedge = 0.0;
  {
    double rp = rpos(s, 0), rm = rpos(s, -1);
    for (int j = 0; j < s->N_p[i]; ++j) {
      double dV = 0.5 * (rp * rp - rm * rm) * s->dphi[j];
      wedge += s->prim[RHO * s->Nz + j] * dV;
    }
  }
  double RMIN = rz[0], RMAX = rz[1], ZMIN = rz[2], ZMAX = rz[3];
  for (int n = 0; n < NS; ++n) scal[n] *= dtout / ((ZMAX - ZMIN) * (RMAX * RMAX - RMIN * RMIN));
  int found = 0;
  double Mdot_req1 = 0.0, r_req1 = 0.0;
  for (int ii = 0; ii < nrg; ++ii) {
    double r = vec[(size_t)ii * NV] / (ZMAX - ZMIN);
    if (r > 1.0 && found == 0) {
      Mdot_req1 = vec[(size_t)ii * NV + 5] * 2. * M_PI * r / (ZMAX - ZMIN);
      r_req1 = r;
      found = 1;
    }
    for (int n = 0; n < NV; ++n) vec[(size_t)ii * NV + n] *= dtout / (ZMAX - ZMIN);
  }
  mdot[0] = Mdot_req1;
  mdot[1] = r_req1;
  mdot[2] = inside1;
  mdot[3] = wedge;
  mdot[4] = s->gm[0].Mdot;
  mdot[5] = s->gm[1].Mdot;
  mdot[6] = s->gm[0].Macc;
  mdot[7] = s->gm[1].Macc;
  mdot[8] = m02[0];
  mdot[9] = m02[1];
  mdot[10] = m04[0];
  mdot[11] = m04[1];
  (void)m01;
}
