// Launchers of the unit-level class-API kernels (unit.hip): the pieces of
// timestep_substep run one call at a time on the current state and the
// unit-level conserved state uc ([NQ][Ntot]).
#pragma once
#include "kernels.h"

enum {
  UNIT_DERIVE = 1,       // uc = prim2cons(prim)             (cell_calc_cons)
  UNIT_COPY,             // RKcons = uc, RKtiph = tiph       (cell_copy)
  UNIT_ADJUST,           // uc = (1-a) uc + a RKcons         (cell_adjust_RK_cons, a = RK)
  UNIT_SRC,              // sources into uc, dt = a; arg: 0 cell_add_src, 1 cell_add_visc_src,
                         // 2 cell_add_visc_src_old, 3 cell_add_split_fictitious
  UNIT_UPDATE_PHI,       // cell_update_phi, RK = a, dt = b
  UNIT_UPDATE_DPHI,      // cell_update_dphi
  UNIT_CALC_PRIM,        // prim = cons2prim(uc)             (cell_calc_prim)
  UNIT_OUTFLOW_Z,        // cell_boundary_outflow_z
};
enum { UNIT_FIXED_R = 1, UNIT_FIXED_Z, UNIT_DAMP };

// one face of riemann_setup_rz / riemann_setup_p: kind 0 r, 1 phi, 2 z; the
// two cells, the annulus of the left cell and the face geometry
struct UnitFace {
  int kind, ann_L;
  long long cL, cR;
  double r, dA, cm, deltaL, deltaR, r_cell_L, r_cell_R;
};

// mdot (UNIT_SRC, arg 0): 2 doubles per zone tile, the tiles' sink accretion
void launch_unit(const DevMesh &m, const LaunchPlan &L, hipStream_t st, int NQ, int op, double *uc, double a,
                 double b, double *cool, double *mdot, int arg);
void launch_unit_bc(const DevMesh &m, const LaunchPlan &L, hipStream_t st, int NQ, const InitRow *rows, int mode,
                    double dt);
void launch_unit_face(const DevMesh &m, hipStream_t st, int NQ, double *uc, const UnitFace &f, double dt);
void launch_unit_add(double *uc, long long x, double v, hipStream_t st);
// both stages of cell_boundary_outflow_r on the current state (kernels.hip)
void launch_bc_outflow_r(const DevMesh &m, const LaunchPlan &L, hipStream_t st, int NQ);
