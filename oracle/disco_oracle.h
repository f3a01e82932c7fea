/*
 * CPU restatement of the Disco2 hydro substep -- TEST INFRASTRUCTURE ONLY.
 *
 * This is the oracle the HIP product is checked against.  It is a plain,
 * sequential C restatement of the reference algorithm that keeps the
 * reference's loop order and floating-point expression order, so that it is
 * bitwise identical to the reference build (gcc x86-64, no FMA).  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Layout: zone(k,i,j) = k*NL + aoff[i] + j with NL = sum_i N_p[i]; every
 * per-zone field is a separate array (q-major for the NQ-vectors).
 */
#ifndef DISCO_ORACLE_H
#define DISCO_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* integer parameter slots */
enum {
  OI_InitialDataType, OI_GravMassType, OI_BoundTypeR, OI_BoundTypeZ, OI_NoInnerBC,
  OI_W_NUMERIC_TYPE, OI_W_ANALYTIC_TYPE, OI_Riemann, OI_VISC_CONST, OI_GRAV2D,
  OI_RhoSinkOn, OI_PHI_ORDER, OI_COOLING, OI_NglobalZ, OI_NgMinR, OI_NgMaxR,
  OI_NgMinZ, OI_NgMaxZ, OI_NUM_C, OI_COUNT
};
/* double parameter slots */
enum {
  OD_GAMMALAW, OD_CFL, OD_PLM, OD_EXPLICIT_VISCOSITY, OD_G_EPS, OD_RHO_FLOOR,
  OD_CS_FLOOR, OD_CS_CAP, OD_VEL_CAP, OD_MassRatio, OD_OrbShrinkTscale,
  OD_OrbShrinkT0, OD_T_MAX, OD_DAMP_TIME, OD_PoRho_r1, OD_RDAMP_INNER, OD_RDAMP_OUTER,
  OD_RMIN, OD_RMAX, OD_COUNT
};

typedef struct OracleSim OracleSim;

OracleSim *oracle_create(int NR, int NZ, int NQ, const int *N_p, const double *r_faces,
                         const double *z_faces, const int *iparams, const double *dparams);
void oracle_destroy(OracleSim *s);
int oracle_nzones(OracleSim *s);
/* initial state; tiph is per zone, prim is q-major [NQ][Nzones] */
void oracle_set_tiph(OracleSim *s, const double *tiph);
void oracle_set_dphi(OracleSim *s, const double *dphi);
void oracle_set_prim(OracleSim *s, const double *prim);
/* the InitialDataType plugin (src/main.c:103-104, after gravMass init) over
 * the current cells; 0, or 17 for an unknown type */
int oracle_initial_data(OracleSim *s);
/* gravMass init + clean_pi + z-periodic sync + calc_cons (src/main.c:103-119) */
void oracle_init(OracleSim *s);
/* one timestep_rk2; returns 0, or a nonzero error code */
int oracle_rk2(OracleSim *s);
int oracle_run(OracleSim *s, int nsteps);
/* timestep_rk2 in pieces (for the conservation ledger): set-up (set_w, dt,
 * cell_copy, gravMass_copy, RK = 0), one substep with its gravMass_move
 * (then RK = 0.5), the end (t += dt, Macc) */
int oracle_step_begin(OracleSim *s);
int oracle_step_substep(OracleSim *s, double fac);
void oracle_step_end(OracleSim *s);
/* the last substep's ledger: the increments dt*dA*F[q] (q < NQ) and
 * dt*dA*Fvisc[SRR, LLL, TAU] that the r faces between the inner ghost annulus
 * and the first interior annulus added to their R cell (NQ + 3 values), the
 * same for the faces between the last interior annulus and the outer ghosts
 * (NQ + 3), and the substep's torque term sum dt dV rho F_phi r (1) */
void oracle_ledger(OracleSim *s, double *out);
/* one timestep_forward_euler (src/TimeStep/timestep_algorithm.c:43-57) */
int oracle_forward_euler(OracleSim *s);
double oracle_torque(OracleSim *s);
/* field: 0 prim, 1 cons, 2 RKcons, 3 grad, 4 gradp (NQ*Nz each);
 *        5 tiph, 6 RKtiph, 7 dphi, 8 wiph (Nz each) */
void oracle_get(OracleSim *s, int field, double *out);
void oracle_put(OracleSim *s, int field, const double *in);
/* t, dt, RK, then 2 x {r, phi, M, omega, RK_r, RK_phi, RK_M, RK_omega, Mdot, Macc} */
void oracle_scalars(OracleSim *s, double *out22);
void oracle_set_scalars(OracleSim *s, const double *in22);
/* unit-level pieces of the substep */
int oracle_face_create(OracleSim *s, int dir);           /* returns face count */
/* faces of the last oracle_face_create(dir): LR[2*nf], geo[6*nf] = {r, dphi, dA,
 * deltaL, deltaR, cm}, offsets[NR] (dir 0) or [NZ] (dir 1) */
void oracle_face_export(OracleSim *s, int dir, int *LR, double *geo, int *offsets);
void oracle_plm_p(OracleSim *s);
void oracle_plm_rz(OracleSim *s, int dir);
double oracle_mindt(OracleSim *s);
void oracle_set_w(OracleSim *s);
int oracle_last_error(OracleSim *s);
/* diagnostics_set (src/Diagnostics/diagnostics_set.c): see disco_oracle.c */
int oracle_diag_neq(OracleSim *s);
void oracle_diag_measure(OracleSim *s, double dtout, const double *rz, double *vec, double *scal,
                         double *equat, double *mdot);

#ifdef __cplusplus
}
#endif
#endif
