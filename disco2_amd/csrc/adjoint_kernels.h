// Device buffers and launchers of the adjoint (adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "adjoint.h"
#include "kernels.h"

// Adjoints of one substep, all q-major like the state.  aPn: of the
// substep's output primitives (ghost entries are consumed by the remap);
// aP: of its input primitives (written); aR: of RKcons (written by the
// second substep, read by the first); au: of each cell's conserved state
// before cons2prim; agrad / agradp: of the r and phi slopes; aSL / aSR: of
// each face's slope S, from its L and its R cell.
struct AdjMesh {
  double *aP, *aPn, *aR, *au, *agrad, *agradp, *aSL, *aSR;
  ScalarAdj *part;     // per-tile partial sums of the scalar adjoints
  double wl;           // weight of the torque term in the loss (1/T)
  double fac, RK;      // the substep's dt factor and RK weight
  int copy_first;      // first substep: RKcons = prim2cons(prim)
  double lam_dt;       // adjoint of the step's dt (CFL backward)
  double *arot;        // per cell: adjoint of its annulus's rotation angle (summed per annulus)
};

void launch_adjoint_substep(const DevMesh &m, const AdjMesh &a, const LaunchPlan &L, hipStream_t st,
                            ScalarAdj *out);
// pick: device int scratch for the index of the cell that sets dt
void launch_adjoint_dt(const DevMesh &m, const AdjMesh &a, const LaunchPlan &L, hipStream_t st,
                       unsigned long long dtbits, int *pick, ScalarAdj *out);
void launch_torque_sum(const double *tiles, int n, double *acc, hipStream_t st);
// per annulus sum of a.arot into out[annulus]
void launch_adjoint_rot(const DevMesh &m, const AdjMesh &a, const LaunchPlan &L, hipStream_t st, double *out);
