// Engine: device-resident state of a batch of disks and the launch sequence
// of timestep_rk2 (src/TimeStep/timestep_algorithm.c:11-41).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "disco_internal.h"
#include "kernels.h"
#include "unit.h"

// What the adjoint keeps of one substep: the state it produced (ghosts
// included) and everything its backward reads.  The forward writes these
// directly (the engine's buffers are pointed at them while it records).
struct AdjSubRec {
  double *prim, *tiph, *dphi;
  double *grad, *gradp, *usave, *fdphi, *fcm;
  double *dt_tile;  // the interior tiles' CFL minima (written in a step's first substep)
  int *fL, *fR, *fo_r, *run;
  PairState *pst;
  signed char *ftype;
};

struct Engine {
  int ndisk = 0, NQ = 5;
  std::vector<Sim *> sims;
  std::vector<DiskParams> hdp;
  std::vector<AnnInfo> hann;
  std::vector<PairInfo> hpairs;
  std::vector<long long> zbase, nzones;
  long long Ntot = 0, NFtot = 0;
  long long nrun = 1;  // run_start slots
  int nann = 0, n_rpairs = 0, n_zpairs = 0;

  // device allocations
  DiskParams *d_dp = nullptr;
  DiskScalars *d_sc = nullptr;
  AnnInfo *d_ann = nullptr;
  PairInfo *d_pairs = nullptr;
  PairState *d_pst = nullptr;
  double *d_rf = nullptr, *d_zf = nullptr;
  double *d_prim[2] = {nullptr, nullptr}, *d_RKcons = nullptr, *d_grad = nullptr, *d_gradp = nullptr,
         *d_grad_r = nullptr;
  double *d_tiph[2] = {nullptr, nullptr};
  double *d_RKtiph = nullptr, *d_dphi[2] = {nullptr, nullptr}, *d_wiph = nullptr;
  int *d_fL = nullptr, *d_fR = nullptr;
  double *d_fdphi = nullptr, *d_fcm = nullptr;
  int *d_fo_r = nullptr, *d_fo_z = nullptr, *d_run = nullptr;
  double *d_Fface = nullptr;
  int *d_err = nullptr;
  double *d_dt_tile = nullptr;
  double *d_mdot_tile = nullptr;  // sink accretion per update tile (when some disk has RhoSinkOn)
  double4 *d_ann_rot = nullptr;
  InitRow *d_init = nullptr;  // fixed r boundaries and damping, per annulus (null when no disk has them)
  int2 *d_tiles = nullptr;
  int cur = 0;  // index of the prim / tiph / dphi buffers holding the current state
  bool w_static = false;  // no disk moves its annuli with the gas (C_RIGID): wiph depends on r alone
  bool w_ready = false;   // wiph has been set (a launch for it is enqueued)
  double *adj_usave = nullptr, *adj_torque = nullptr;  // adjoint recording (mesh() passes them on)
  double *adj_torque_acc = nullptr;
  DiskScalars *adj_sc_rec = nullptr;
  signed char *adj_ftype = nullptr;
  void *adj_mem = nullptr;  // the adjoint's working memory, kept for later calls
  size_t adj_cap = 0;

  // launch schedule (the autotuner's choice; the results do not depend on it):
  // XCD-banded tiles, the face builder's pairs largest first, the outer
  // annulus staged in LDS
  // fuse: bit 0 / 1 evaluate the r / z faces' increments inside the update
  // instead of storing them (each face is then solved by both of its cells)
  // zb: 3D disks' zone, update and z-face tiles in blocks of 2^zb layers,
  // annulus column by annulus column inside a block (0: table order)
  struct Schedule {
    int xcd = 1, largest_first = 0, stage = 1, fuse = 0, zb = 0, fw = 0;
    int code() const { return xcd | largest_first << 1 | stage << 2 | fuse << 3 | zb << 5 | fw << 8; }
    static Schedule of(int k) {
      Schedule s;
      s.xcd = k & 1;
      s.largest_first = (k >> 1) & 1;
      s.stage = (k >> 2) & 1;
      s.fuse = (k >> 3) & 3;
      s.zb = (k >> 5) & 7;
      s.fw = (k >> 8) & 3;
      return s;
    }
    int face_threads() const { return 1024 >> (fw < 3 ? fw : 0); }
  } sched;
  std::vector<int2> h_ztiles, h_utiles, h_zft;       // table-order lists
  int2 *d_tiles_zb[8] = {};                          // [zone | update | z face] tiles per layer block
  const int2 *tl_zb[8][3] = {};                      // (zb = 0: slices of d_tiles)
  bool build_zblock(int zb);
  bool schedule_distinct(int k) const;
  int stage_r_max = 0, stage_z_max = 0;  // LDS staging the plan allows
  int *d_pair_perm = nullptr;
  void set_schedule(const Schedule &s);
  // times the candidate schedules on a scratch copy of the batch and keeps
  // the fastest; ms[k] per candidate k (Schedule::code; -1: not timed)
  int autotune(int nsteps, double *ms, int nms);

  LaunchPlan plan{};
  hipStream_t st = nullptr;
  // the adjoint's recompute stream (default priority): a segment's forward
  // recompute runs on it while the reverse sweep of the next segment runs on st
  hipStream_t st_aux = nullptr;
  // the adjoint's second stream and its events (recompute done, sweep done
  // per record, loss pass done), made by the first adjoint call
  hipEvent_t adj_ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_cur = -1;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  bool ok = false;

  Engine(int n, Sim **sims);
  ~Engine();
  DevMesh mesh() const;
  void upload(int disk, int field, const double *src);
  void download(int disk, int field, double *dst);
  void put_scalars(int disk, const DiskScalars &s);
  DiskScalars get_scalars(int disk);
  void init_state();                 // gravMass init, z sync of prim/tiph, scalars
  void enqueue_rk2();
  void enqueue_set_w();
  void enqueue_euler();
  void enqueue_substep(double fac, int copy_first, int recompute_w, int pre_ops, int post_ops);
  void run(int nsteps, bool use_graph);
  void synchronize();
  int check_error();                 // 0, or sets the sticky library error

  // adjoint (adjoint_host.cpp)
  bool adjoint_scope();
  // one RK2 step; with rec (two substep records) and screc (three scalar
  // records: as k_mindt saw them, at the start of substep 1 and 2) it records
  void adjoint_step(const AdjSubRec *rec, DiskScalars *screc, double *torque, double *acc);
  void adjoint_redirect(const AdjSubRec &r);  // the next substep writes into r
  void adjoint_restore(const double *src, bool rk);
  int torque_loss(int nsteps, double *loss);
  // torque_loss with its accumulator and tile sums in the library's fixed
  // device scratch: no device allocation (adjoint_host.cpp)
  int torque_loss_noalloc(int nsteps, double *loss);

  // diagnostics (diagnostics.cpp): one measurement of diagnostics_set on disk
  // d (src/Diagnostics/diagnostics_set.c:13-303): vec [NumR][17], scal[16],
  // equat [N_eq][18], mdot[12] as the DiagMdot.dat line after its t
  int diag_sizes(int d, int *nrg, int *neq) const;
  // (equat may be null: the cells stay on the device for diag_equat)
  int diag_measure(int d, double dtout, const double *rz, double *vec, double *scal, double *equat, double *mdot);
  int diag_equat(double *out);  // the last measurement's equatorial cells

  // checkpoints (checkpoint.cpp): disk d's cell rows {tiph, r, z, prim}
  // (src/IO/io_buffer.c), nzones[d] x (NQ + 3) doubles on the host
  int checkpoint_pack(int d, double *rows);
  int checkpoint_unpack(int d, const double *rows);
  int checkpoint_write(int d, const char *fname);
  int checkpoint_read(int d, const char *fname);
  // unit-level class API (capi.cpp, unit.hip): the reference's cell->cons as
  // its own array while a driver calls the substep's pieces one by one; it
  // follows the primitives (prim2cons) until a unit-level call changes it
  double *d_ucons = nullptr, *d_cool = nullptr, *d_umdot = nullptr;
  InitRow *d_urows = nullptr;
  bool ucons_live = false;
  double *unit_cons();      // the live conserved state (allocated and derived on first use); null on error
  // conservation ledger (2D disks): with d_led_torque set, k_update stores
  // its tiles' torque terms; ledger_last reads what the last substep's
  // boundary r faces and torque added (layout of disco_cell_ledger)
  double *d_led_torque = nullptr;
  int ledger_enable(bool on);
  int ledger_last(int disk, double *out);
  InitRow *unit_rows();     // the disk's single-cell plugin rows on the device; null on error
  void *ckpt_mem = nullptr;
  size_t ckpt_cap = 0;
  void *diag_mem = nullptr;
  size_t diag_cap = 0;
  double *diag_equat_dev = nullptr;
  size_t diag_equat_n = 0;
  int adjoint(int nsteps, int every, double *loss, double *g_alpha, double *g_q, double *g_prim0);
};

bool hip_check(hipError_t e, const char *what);
