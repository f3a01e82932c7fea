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
// each face's slope S, from its L and its R cell; fr: per r face, what its
// flux and slope backward hand its two cells (AFR_* slots of NFtot each).
enum {
  AFR_HL = 0,    // Riemann part: adjoint of primL (5)
  AFR_HR = 5,    // and of primR (5)
  AFR_WL = 10,   // viscous part: adjoint of W_L (5)
  AFR_WR = 15,   // and of W_R (5)
  AFR_G = 20,    // viscous part: to both cells' phi slopes of v_r, omega (2)
  AFR_ROT = 22,  // flux parts: to the L and the R cell's annulus rotation (2)
  AFR_WB = 24,   // r slope sums: adjoint of the face width from the L and the R cell (2)
  AFR_N = 26
};
struct AdjMesh {
  double *aP, *aPn, *aR, *au, *agrad, *agradp, *aSL, *aSR, *fr;
  double *part;        // per-tile partial sums of the scalar adjoints: 9 rows of npart
  long long npart;
  double wl;           // weight of the torque term in the loss (1/T)
  double fac;          // the substep's dt factor (its RK weight is m.sc's)
  int copy_first;      // first substep: RKcons = prim2cons(prim)
  const double *lam_dt;  // device: adjoint of the step's dt (CFL backward)
  double *arot;        // per cell: adjoint of its annulus's rotation angle (summed per annulus)
  int visc;            // the disk is viscous: launch the viscous face backward kernels
};

// reverse-sweep sums of the adjoints of the launch-uniform scalars: the
// masses' phases and radii, their masses, alpha; lam_dt of the current step
struct AdjChain {
  double Lphi[2], Lr[2], LM[2], Lalpha, lam_dt;
};

// the backward kernels of one substep; their per-tile partial sums of the
// scalar adjoints go to a.part (launch_adjoint_reduce sums them)
void launch_adjoint_substep(const DevMesh &m, const AdjMesh &a, const LaunchPlan &L, hipStream_t st);
// out[b] = the sum of the substep partials in a.part + b * 9 * a.npart, for
// b < nbuf, one workgroup each
void launch_adjoint_reduce(const AdjMesh &a, const LaunchPlan &L, hipStream_t st, int nbuf, ScalarAdj *out);
// pick: device int scratch (0x7fffffff: none yet) that receives the lowest
// interior cell whose CFL step is the step's minimum dtbits (device)
void launch_adjoint_pick(const DevMesh &m, const LaunchPlan &L, hipStream_t st, const unsigned long long *dtbits,
                         int *pick);
// per annulus sum of a.arot into out[annulus]; clears a.arot
void launch_adjoint_rot(const DevMesh &m, const AdjMesh &a, const LaunchPlan &L, hipStream_t st, double *out);
void launch_copy_scalars(DiskScalars *dst, const DiskScalars *src, hipStream_t st);
// lam_dt of a step from its substeps' scalar adjoints S1, S2, the second
// substep's annulus rotation adjoints th2 and the sums so far (m: annuli,
// wiph, and the scalars the forward's CFL minimum saw); then the CFL step's backward at the
// picked cell into a.aP and a.arot, its scalar adjoints into S0, and pick
// reset
void launch_adjoint_chain_dt(const DevMesh &m, const AdjMesh &a, int nann, const DiskScalars *A,
                             const DiskScalars *B, const ScalarAdj *S1, const ScalarAdj *S2, const double *th2,
                             const double *Ltheta, double lamT, AdjChain *cs, int *pick, ScalarAdj *S0,
                             hipStream_t st);
// adds a step's contributions (S0: the CFL step's) to the sums
void launch_adjoint_chain_acc(int nann, const DiskScalars *pre, const DiskScalars *A, const DiskScalars *B,
                              const ScalarAdj *S0, const ScalarAdj *S1, const ScalarAdj *S2, const double *th1,
                              const double *th2, double *Ltheta, AdjChain *cs, hipStream_t st);

// Fixed device scratch of the allocation-free torque loss: the accumulator
// and one slot per k_update tile (512 KiB: batches of up to 65,535 update
// tiles, about 16.7M zones), on up to 64 GPUs per process.
#define TORQUE_SCRATCH_DOUBLES (1 << 16)
#define TORQUE_SCRATCH_MAX_DEVICES 64
// this GPU's scratch address (null on error); p[0, n) = 0 on the stream
double *torque_scratch_device();
void launch_zero(double *p, int n, hipStream_t st);
