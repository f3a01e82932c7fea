// Device-side view of a batch and the host-side launchers of kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "disco_types.h"

struct DevMesh {
  const DiskParams *dp;
  DiskScalars *sc;
  const AnnInfo *ann;
  const PairInfo *pairs;
  PairState *pst;
  const double *rf, *zf;  // face-position pools (per-disk offsets geo_r / geo_z)
  double *prim, *RKcons, *grad, *gradp;
  double *grad_r;         // 3D: the r slopes, beside the z slopes in grad (null in 2D)
  double *prim_new;       // primitives after the substep (k_update writes them)
  const double *tiph;     // cell edges at the start of the substep
  double *tiph_new;       // cell edges after the substep
  double *RKtiph, *dphi;
  double *dphi_new;       // cell widths after the substep
  double *wiph;           // per annulus
  int *fL, *fR;
  double *fdphi, *fcm;
  int *fo_r, *fo_z;       // per zone: first face of the cell as inner cell of its r / z pair
  int *run_start;
  double *Fface;          // [NQ+3][NFtot] increments of the r/z overlap faces
  double *dt_tile;        // CFL minimum of each interior tile (k_mindt), reduced per disk by k_scalar
  double *mdot_tile;      // sink: per update tile, the two masses' accretion (2 per tile; null without sinks)
  double4 *ann_rot;       // per annulus: rOm_a, rdrOm_a, dtOm_a at the centre, rdrOm_a at the outer face
  const InitRow *init;    // per annulus: fixed r boundaries and damping (null when unused)
  // adjoint recording (null in plain runs): k_update stores the conserved
  // state it converts (u_save, NQ x Ntot) and its tiles' torque sums
  double *u_save, *torque_tile;
  // and, in the launches that run a disk's scalar operations: torque_acc[d]
  // accumulates the sum of its update tiles' torque (after each update), and
  // sc_rec[3 d ...] receives the scalars of an RK2 step: before and after the
  // step's opening operations (SC_STEP_BEGIN), after the first substep's
  double *torque_acc;
  DiskScalars *sc_rec;
  signed char *ftype;     // the face builder records each face's bounding edges here
  long long Ntot, NFtot;
  int *err;
  // launch schedule (chosen by the batch's autotuner): workgroup -> tile in
  // XCD bands or in launch order; face-builder pairs in table order or the
  // permutation pair_perm (largest outer annulus first)
  int sched_xcd;
  const int *pair_perm;
};

// which edges bound an overlap face, left then right (I: inner annulus,
// O: outer): the face's width and centre move with the two annuli's rotations
enum { FACE_II = 0, FACE_IO = 1, FACE_OO = 2, FACE_OI = 3 };

// LDS the face builder may use to stage an outer annulus (two doubles per
// cell), below the 160 KiB of a gfx950 workgroup minus its static arrays
#define FACE_STAGE_BYTES (144 * 1024)

// k_scalar runs one workgroup of SCALAR_THREADS per disk: all of them reduce
// the disk's tile minima (SC_DT_FOLD), thread 0 does the scalar updates.
#define SCALAR_THREADS 1024

struct LaunchPlan {
  int ndisk, nann;
  int n_rpairs, n_zpairs;
  const int2 *ztiles; int n_ztiles;     // every zone
  const int2 *utiles; int n_utiles;     // every zone, DISCO_TILE - 1 cells per tile (k_update)
  const int2 *stiles; int n_stiles;     // every zone, DISCO_TILE - 2 cells per tile (a phi neighbour each side)
  const int2 *itiles; int n_itiles;     // interior zones (CFL)
  const int2 *rftiles; int n_rftiles;   // r faces
  const int2 *zftiles; int n_zftiles;   // z faces
  int stage_r, stage_z;                 // face builder: outer-annulus cells staged in LDS (0: none)
  int any_visc;                         // some disk has viscosity
  int feat;                             // physics paths its disks take (F_* bits)
  int fuse;                             // schedule: r (1) / z (2) face increments evaluated inside k_update
  int face_threads;                     // schedule: threads per face-builder workgroup (1024, 512, 256)
  const int2 *bc0; int n_bc0;           // ghost annuli 0 and NR-2
  const int2 *bc1; int n_bc1;           // ghost annulus NR-1
  const int2 *sz; int n_sz;             // z ghost layers
  const int2 *binit; int n_binit;       // fixed-boundary ghost annuli and damped annuli
};

// Optional per-kernel timing: when g_kprof is set, every launch is bracketed
// by HIP events on its stream (eager launches only, never under capture).
struct KMark {
  const char *name;
  hipEvent_t a, b;
};
struct KProf {
  KMark *marks;
  int n, cap;
};
extern KProf *g_kprof;
#define KLAUNCH(st, nm, ...)                                              \
  do {                                                                    \
    if (g_kprof && g_kprof->n < g_kprof->cap) {                           \
      KMark &mk_ = g_kprof->marks[g_kprof->n++];                          \
      mk_.name = nm;                                                      \
      hipEventRecord(mk_.a, st);                                          \
      __VA_ARGS__;                                                        \
      hipEventRecord(mk_.b, st);                                          \
    } else {                                                              \
      __VA_ARGS__;                                                        \
    }                                                                     \
  } while (0)

// per-disk scalar operations (bit flags; one launch runs the set in the
// order listed)
enum {
  SC_STEP_BEGIN = 1,    // dt from the CFL minimum, gravMass_copy, RK = 0
  SC_AFTER_UPDATE = 2,  // Mdot, gravMass_update_RK
  SC_SUB1_END = 4,      // gravMass_move(t, dt), RK = 0.5
  SC_SUB2_END = 8,      // gravMass_move(t, dt/2), t += dt, Macc
  SC_EULER_END = 16,    // gravMass_move(t, dt), t += dt
  SC_SUB_BEGIN = 32,    // gravMass_clean_pi, Mdot accumulators
  SC_RESET_DT = 64,
  SC_MASS = 128,        // launch-uniform gravity terms
  SC_DT_FOLD = 256,     // CFL minimum from the tile minima (runs first; STEP_BEGIN implies it)
  SC_PLM_DT = 512,      // pre_ops of launch_substep: the phi slopes take the CFL tile minima, and
                        // the pre-substep scalar launch (with its fold) follows them
};

void launch_substep(const DevMesh &m, const LaunchPlan &L, hipStream_t st, int NQ, double fac, int copy_first,
                    int recompute_w, int pre_ops, int post_ops);
bool prepare_launches(const LaunchPlan &L);  // kernel attributes the plan needs
// the compiled variant (F_* set) that launch_substep uses for these needs
int feature_variant(int need, int NQ);
void launch_scalar(const DevMesh &m, const LaunchPlan &L, hipStream_t st, int op);
void launch_set_w(const DevMesh &m, const LaunchPlan &L, hipStream_t st);
void launch_mindt(const DevMesh &m, const LaunchPlan &L, hipStream_t st, int NQ);
void launch_faces(const DevMesh &m, const LaunchPlan &L, hipStream_t st, int NQ, int dir);
void launch_plm(const DevMesh &m, const LaunchPlan &L, hipStream_t st, int NQ, int which);
void launch_copy(double *dst, const double *src, long long n, hipStream_t st);

// checkpoint rows (checkpoint.hip): nann = the disk's annuli (NR x NZ);
// unpack writes the primitives and tiph (the current state's buffers)
void launch_ckpt_pack(const DevMesh &m, int disk, int nann, double *out, hipStream_t st);
void launch_ckpt_unpack(const DevMesh &m, int disk, int nann, const double *in, double *tiph, hipStream_t st);

// NUM_Q = 5 + NUM_N, NUM_N = 0..3: one instantiation of every kernel each
#define DISCO_NQ_DISPATCH(NQ, CALL)      \
  switch (NQ) {                          \
    case 5: { constexpr int Q = 5; CALL; } break; \
    case 6: { constexpr int Q = 6; CALL; } break; \
    case 7: { constexpr int Q = 7; CALL; } break; \
    default: { constexpr int Q = 8; CALL; } break; \
  }
