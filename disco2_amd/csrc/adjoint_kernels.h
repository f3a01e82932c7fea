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

// Work that rides in the workgroups past the tiles of a substep's backward
// launches instead of launches of its own: the per-annulus sums of a
// rotation-adjoint buffer (rot_src non-null) and the step's CFL cell (pick
// non-null), both in the substep's update launch, and the reduction
// of both substeps' scalar partials (red non-null, in its slope launch).
// The two substeps keep their cells' rotation adjoints in two buffers, so
// that one substep's sums run while the other writes.
#define ADJ_RED_BLOCKS 32  // workgroups per partial buffer in the reduction
struct AdjExtra {
  double *rot_src = nullptr;  // per cell rotation adjoints, summed per annulus into rot_out and cleared
  double *rot_out = nullptr;  // [nann]
  const double *red_src = nullptr;  // the second substep's partial buffer, the first's after it
  double *red = nullptr;            // [2][ADJ_RED_BLOCKS][9]: their partial sums
  // the CFL cell: the step-start state, the scalars its CFL minimum saw, the
  // interior tiles' minima and the minimum; *pick receives the lowest
  // interior zone whose step is the minimum and the gradient of its step
  const double *pick_prim = nullptr, *pick_tiph = nullptr, *pick_dphi = nullptr, *pick_dt_tile = nullptr;
  const DiskScalars *pick_sc = nullptr;
  const unsigned long long *dtbits = nullptr;
  struct CflGrad *pick = nullptr;
};
// The cell whose CFL step is the step's dt (c = INT_MAX: none), its annulus,
// and the gradient of its step: on its primitives (prb), its annulus's
// rotation (tcb) and the scalars (sa: the ScalarAdj fields in order).
struct CflGrad {
  int c, ann;
  double prb[5], tcb, sa[9];
};

// the backward kernels of one substep and the work of x; their per-tile
// partial sums of the scalar adjoints go to a.part
void launch_adjoint_substep(const DevMesh &m, const AdjMesh &a, const LaunchPlan &L, hipStream_t st,
                            const AdjExtra &x);
void launch_copy_scalars(DiskScalars *dst, const DiskScalars *src, hipStream_t st);
// The scalar chain of one step, one workgroup: the two substeps' scalar
// adjoints S2, S1 from red; lam_dt of the step from them, the second
// substep's annulus rotation adjoints th2, the first substep's of the step
// after this one (th1n: those are summed one step late) and the sums so far
// (m: annuli, wiph, the step-start state and the scalars the forward's CFL
// minimum saw); the CFL step's backward, lam_dt times *pick's gradient,
// into a.aP (its rotation adjoint into Ltheta of its annulus, its scalar
// adjoints S0); the step's contributions added to the sums (Ltheta += th1n +
// th2, cs).
void launch_adjoint_chain(const DevMesh &m, const AdjMesh &a, const CflGrad *pick, int nann, const DiskScalars *pre,
                          const DiskScalars *A, const DiskScalars *B, const double *red, const double *th1n,
                          const double *th2, double *Ltheta, double lamT, AdjChain *cs, hipStream_t st);

// Fixed device scratch of the allocation-free torque loss: the accumulator
// and one slot per k_update tile (512 KiB: batches of up to 65,535 update
// tiles, about 16.7M zones), on up to 64 GPUs per process.
#define TORQUE_SCRATCH_DOUBLES (1 << 16)
#define TORQUE_SCRATCH_MAX_DEVICES 64
// this GPU's scratch address (null on error); p[0, n) = 0 on the stream
double *torque_scratch_device();
void launch_zero(double *p, int n, hipStream_t st);
