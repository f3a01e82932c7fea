// Host side of the adjoint: the loss, the checkpointed forward, the
// recompute of each substep and the reverse sweep, and the chain through
// the launch-uniform scalars (dt, the binary's phase, its masses and radii).
//
// Loss (the time-averaged gravitational torque on the gas):
//   L = (1/T) sum_steps sum_substeps sum_interior dt_s dV rho F_phi r,
// with dt_s the substep's dt (dt, then dt/2), F_phi the phi force of the two
// masses at the cell centre (the LLL source of cell_source.c:171) and T the
// simulated time.  Gradients: d L / d alpha (Explicit_Viscosity),
// d L / d q (MassRatio) and d L / d prim0 (every zone's initial primitives).
//
// Every step's dt depends on the state through its CFL minimum; the binary's
// phase advances by omega * dt and every annulus turns by (w / r) * dt, which
// moves the cells against the masses and the overlap faces against each
// other: all three chains are carried (per-annulus rotation adjoints).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "adjoint_kernels.h"
#include "engine.h"

namespace {

template <class T>
bool dalloc0(T **p, size_t n, const char *what) {
  if (!hip_check(hipMalloc((void **)p, (n ? n : 1) * sizeof(T)), what)) return false;
  return hip_check(hipMemset(*p, 0, (n ? n : 1) * sizeof(T)), what);
}

}  // namespace

bool Engine::adjoint_scope() {
  const DiskParams &P = hdp[0];
  const char *why = nullptr;
  if (ndisk != 1)
    why = "one disk per batch";
  else if (NQ != 5)
    why = "NUM_Q = 5";
  else if (P.NglobalZ != 1)
    why = "2D (NumZ = 1)";
  else if (P.w_analytic != A_FIXED)
    why = "w_analytic_type 0";
  else if (!(P.w_numeric == C_FIXED || P.w_numeric == C_KEPLER || P.w_numeric == C_OMEGA20 ||
             P.w_numeric == C_POWERLAW))
    why = "a mesh velocity independent of the state (w_numeric_type 0, 3, 4 or 6)";
  else if (P.sink == 1 || P.cooling == 1)
    why = "no sink and no cooling";
  else if (P.visc > 0.0 && P.visc_const != 1 && P.initial_data_type == 0)
    why = "constant or alpha-law viscosity";
  else if (P.bound_r != BOUND_OUTFLOW)
    why = "outflow r boundaries";
  else if (plan.n_binit > 0)
    why = "no boundary damping";
  if (why) {
    disco_set_error(12, std::string("the adjoint supports ") + why);
    return false;
  }
  return true;
}

void Engine::adjoint_redirect(const AdjSubRec &r) {
  d_prim[1 - cur] = r.prim;
  d_tiph[1 - cur] = r.tiph;
  d_dphi[1 - cur] = r.dphi;
  d_grad = r.grad;
  d_gradp = r.gradp;
  d_fL = r.fL;
  d_fR = r.fR;
  d_fdphi = r.fdphi;
  d_fcm = r.fcm;
  d_fo_r = r.fo_r;
  d_run = r.run;
  d_pst = r.pst;
  d_dt_tile = r.dt_tile;
  adj_usave = r.usave;  // k_update records what cons2prim converts
  adj_ftype = r.ftype;  // the face builder records the face types
}

// One RK2 step as enqueue_rk2 launches it, recording: the scalars (the
// scalar launches write screc[0..2]: before and after the step's opening
// operations, after the first substep's), the update tiles' torque, added up
// into acc by the substeps' scalar launches, and (rec) the substeps' state,
// slopes, faces and converted conserved states.
void Engine::adjoint_step(const AdjSubRec *rec, DiskScalars *screc, double *torque, double *acc) {
  enqueue_set_w();
  adj_sc_rec = screc;
  adj_torque = torque;
  adj_torque_acc = acc;
  if (rec) adjoint_redirect(rec[0]);
  enqueue_substep(1.0, 1, 0, SC_PLM_DT | SC_STEP_BEGIN | SC_SUB_BEGIN, SC_AFTER_UPDATE | SC_SUB1_END | SC_SUB_BEGIN);
  if (rec) adjoint_redirect(rec[1]);
  enqueue_substep(0.5, 0, w_static ? 0 : 1, 0, SC_AFTER_UPDATE | SC_SUB2_END);
  adj_sc_rec = nullptr;
  adj_torque = nullptr;
  adj_torque_acc = nullptr;
  adj_usave = nullptr;
  adj_ftype = nullptr;
}

void Engine::adjoint_restore(const double *src, bool rk) {
  const size_t nz = (size_t)Ntot, nq = (size_t)NQ * Ntot;
  hipMemcpyAsync(d_prim[cur], src, nq * 8, hipMemcpyDeviceToDevice, st);
  hipMemcpyAsync(d_tiph[cur], src + nq, nz * 8, hipMemcpyDeviceToDevice, st);
  hipMemcpyAsync(d_dphi[cur], src + nq + nz, nz * 8, hipMemcpyDeviceToDevice, st);
  hipMemcpyAsync(d_RKtiph, src + nq + 2 * nz, nz * 8, hipMemcpyDeviceToDevice, st);
  if (rk) hipMemcpyAsync(d_RKcons, src + nq + 3 * nz, nq * 8, hipMemcpyDeviceToDevice, st);
}

int Engine::torque_loss(int nsteps, double *loss) {
  if (!ok || !adjoint_scope()) return -1;
  double *torque = nullptr, *acc = nullptr;
  if (!dalloc0(&torque, (size_t)plan.n_utiles, "torque") || !dalloc0(&acc, 1, "torque sum")) return -1;
  const double t0 = get_scalars(0).t;
  for (int n = 0; n < nsteps; ++n) adjoint_step(nullptr, nullptr, torque, acc);
  const double T = get_scalars(0).t - t0;
  double I = 0.0;
  hip_check(hipMemcpy(&I, acc, 8, hipMemcpyDeviceToHost), "torque D2H");
  hipFree(torque);
  hipFree(acc);
  *loss = T > 0.0 ? I / T : 0.0;
  return check_error() ? -1 : 0;
}

// torque_loss without device allocation: slot 0 of the fixed scratch is the
// accumulator (cleared by a fill kernel), slots 1.. the tile sums, which
// every k_update tile overwrites before the substep's scalar launch adds
// them up.
int Engine::torque_loss_noalloc(int nsteps, double *loss) {
  if (!ok || !adjoint_scope()) return -1;
  if ((long long)plan.n_utiles + 1 > TORQUE_SCRATCH_DOUBLES) {
    disco_set_error(16, "torque_loss_noalloc: " + std::to_string(plan.n_utiles) +
                            " update tiles exceed the fixed scratch of " +
                            std::to_string(TORQUE_SCRATCH_DOUBLES - 1));
    return -1;
  }
  double *scr = torque_scratch_device();
  if (!scr) {
    disco_set_error(16, "torque_loss_noalloc: no device address for the fixed scratch");
    return -1;
  }
  double *acc = scr, *torque = scr + 1;
  launch_zero(acc, 1, st);
  const double t0 = get_scalars(0).t;
  for (int n = 0; n < nsteps; ++n) adjoint_step(nullptr, nullptr, torque, acc);
  const double T = get_scalars(0).t - t0;
  double I = 0.0;
  hip_check(hipMemcpy(&I, acc, 8, hipMemcpyDeviceToHost), "torque D2H");
  *loss = T > 0.0 ? I / T : 0.0;
  return check_error() ? -1 : 0;
}

namespace {

// device buffers of one adjoint run, carved from one allocation
struct Arena {
  char *base = nullptr;
  size_t off = 0;
  template <class T>
  T *take(size_t n) {
    off = (off + 255) & ~(size_t)255;
    T *p = (T *)((uintptr_t)base + off);
    off += (n ? n : 1) * sizeof(T);
    return p;
  }
};

}  // namespace

int Engine::adjoint(int nsteps, int every, double *loss, double *g_alpha, double *g_q, double *g_prim0) {
  if (!ok || !adjoint_scope()) return -1;
  if (nsteps < 1) {
    disco_set_error(12, "adjoint: nsteps must be >= 1");
    return -1;
  }
  if (every < 1 || every > nsteps) every = nsteps;
  // the second stream and the events of the checkpointed sweep, made once
  // (by the first call, whatever its interval) and kept like the working
  // memory
  if (!st_aux && !hip_check(hipStreamCreate(&st_aux), "adjoint stream")) return -1;
  for (hipEvent_t &e : adj_ev)
    if (!e && !hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "adjoint event")) return -1;
  const size_t nz = (size_t)Ntot, nq = (size_t)NQ * Ntot, nf = (size_t)NFtot;
  const size_t wA = nq + 3 * nz;  // doubles per checkpoint: prim, tiph, dphi, RKtiph
  const int nseg = (nsteps + every - 1) / every;
  const int nz_t = plan.n_ztiles, nu_t = plan.n_utiles, nrf_t = plan.n_rftiles, ni_t = plan.n_itiles;
  const size_t npart = std::max((size_t)(nz_t + 2 * nu_t + 2 * nrf_t), (size_t)ni_t);
  const size_t npairs = hpairs.size();

  // one allocation: the adjoint fields, the checkpoints, the record of one
  // segment (2 * every substeps) -- two records with several segments, so
  // that the recompute of one segment (on st_aux) overlaps the reverse sweep
  // of the next (on st); laid out once to size it, once to carve it
  const int nbuf = nseg > 1 ? 2 : 1;
  double *aP, *aPn, *aR, *au, *agrad, *agradp, *aSL, *aSR, *fr, *aFS, *arot2, *arot1, *torque, *acc, *ckpt, *fin, *th1,
      *th2, *Ltheta;
  DiskScalars *screcs[2] = {nullptr, nullptr}, *ckpt_sc, *fin_sc;
  double *part;
  double *red;
  CflGrad *pick;
  AdjChain *cs;
  std::vector<AdjSubRec> recs[2] = {std::vector<AdjSubRec>(2 * every), std::vector<AdjSubRec>(nbuf > 1 ? 2 * every : 0)};
  Arena ar;
  size_t zero_bytes = 0;
  auto layout = [&]() {
    ar.off = 0;
    // first the fields that start at zero (one memset)
    aPn = ar.take<double>(nq);
    arot2 = ar.take<double>(nz);  // the cells' rotation adjoints of the second substep
    arot1 = ar.take<double>(nz);  // and of the first
    acc = ar.take<double>(1);
    Ltheta = ar.take<double>((size_t)nann);
    red = ar.take<double>(2 * ADJ_RED_BLOCKS * 9);  // the two substeps' scalar partials, reduced per block
    pick = ar.take<CflGrad>(1);
    cs = ar.take<AdjChain>(1);
    zero_bytes = ar.off;
    aP = ar.take<double>(nq);
    aR = ar.take<double>(nq);
    au = ar.take<double>(nq);
    agrad = ar.take<double>(nq);
    agradp = ar.take<double>(nq);
    aSL = ar.take<double>(NQ * nf);
    aSR = ar.take<double>(NQ * nf);
    fr = ar.take<double>((size_t)AFR_N * nf);
    aFS = ar.take<double>((size_t)NQ * nf);  // the reverse sweep's face slopes (the forward's Fface is in use)
    torque = ar.take<double>((size_t)plan.n_utiles);
    ckpt = ar.take<double>(wA * nseg);
    fin = ar.take<double>(nseg > 1 ? wA : 0);
    th1 = ar.take<double>((size_t)nann);
    th2 = ar.take<double>((size_t)nann);
    for (int b = 0; b < nbuf; ++b) screcs[b] = ar.take<DiskScalars>((size_t)3 * every);
    ckpt_sc = ar.take<DiskScalars>((size_t)nseg);
    fin_sc = ar.take<DiskScalars>(1);
    part = ar.take<double>(2 * 9 * npart);  // the two substeps' partial sums
    for (int b = 0; b < nbuf; ++b) {
      for (auto &r : recs[b]) {
        r.prim = ar.take<double>(nq);
        r.grad = ar.take<double>(nq);
        r.gradp = ar.take<double>(nq);
        r.usave = ar.take<double>(nq);
        r.tiph = ar.take<double>(nz);
        r.dphi = ar.take<double>(nz);
        r.fdphi = ar.take<double>(nf);
        r.fcm = ar.take<double>(nf);
        r.fL = ar.take<int>(nf);
        r.fR = ar.take<int>(nf);
        r.fo_r = ar.take<int>(nz);
        r.run = ar.take<int>((size_t)nrun);
        r.pst = ar.take<PairState>(npairs);
        r.ftype = ar.take<signed char>(nf);
        r.dt_tile = ar.take<double>((size_t)std::max(ni_t, 1));
      }
    }
  };
  layout();
  if (ar.off > adj_cap) {  // the working memory is kept for later calls
    if (adj_mem) hipFree(adj_mem);
    adj_mem = nullptr;
    adj_cap = 0;
    if (hipMalloc(&adj_mem, ar.off) != hipSuccess) {
      (void)hipGetLastError();
      adj_mem = nullptr;
      disco_set_error(13, "adjoint: the record of a segment (" + std::to_string(ar.off >> 20) +
                              " MiB) does not fit in device memory; lower checkpoint_every");
      return -1;
    }
    adj_cap = ar.off;
  }
  ar.base = (char *)adj_mem;
  layout();
  if (!hip_check(hipMemsetAsync(ar.base, 0, zero_bytes, st), "adjoint")) return -1;

  // the engine's own buffers, pointed back at after each recorded segment
  struct Own {
    double *prim[2], *tiph[2], *dphi[2], *grad, *gradp, *fdphi, *fcm, *dt_tile;
    int *fL, *fR, *fo_r, *run;
    PairState *pst;
  } own = {{d_prim[0], d_prim[1]}, {d_tiph[0], d_tiph[1]}, {d_dphi[0], d_dphi[1]}, d_grad, d_gradp, d_fdphi, d_fcm,
           d_dt_tile, d_fL, d_fR, d_fo_r, d_run, d_pst};
  auto end_record = [&]() {
    hipMemcpyAsync(own.prim[cur], d_prim[cur], nq * 8, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(own.tiph[cur], d_tiph[cur], nz * 8, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(own.dphi[cur], d_dphi[cur], nz * 8, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(own.grad, d_grad, nq * 8, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(own.gradp, d_gradp, nq * 8, hipMemcpyDeviceToDevice, st);
    for (int b = 0; b < 2; ++b) {
      d_prim[b] = own.prim[b];
      d_tiph[b] = own.tiph[b];
      d_dphi[b] = own.dphi[b];
    }
    d_grad = own.grad;
    d_gradp = own.gradp;
    d_fdphi = own.fdphi;
    d_fcm = own.fcm;
    d_fL = own.fL;
    d_fR = own.fR;
    d_fo_r = own.fo_r;
    d_run = own.run;
    d_pst = own.pst;
    d_dt_tile = own.dt_tile;
  };
  auto snapshot = [&](double *c) {
    hipMemcpyAsync(c, d_prim[cur], nq * 8, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(c + nq, d_tiph[cur], nz * 8, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(c + nq + nz, d_dphi[cur], nz * 8, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(c + nq + 2 * nz, d_RKtiph, nz * 8, hipMemcpyDeviceToDevice, st);
  };

  // the record of segment seg: the last in record 0, then alternating
  auto buf_of = [&](int seg) { return nbuf > 1 ? (nseg - 1 - seg) & 1 : 0; };

  // forward: the loss, a checkpoint at the start of every segment, and the
  // record of the last nbuf segments (one per record)
  const int rec0 = (nseg - nbuf) * every;
  const double t0 = get_scalars(0).t;
  for (int n = 0; n < nsteps; ++n) {
    if (n % every == 0) {
      snapshot(ckpt + wA * (n / every));
      launch_copy_scalars(ckpt_sc + n / every, d_sc, st);
    }
    if (n >= rec0) {
      const int b = buf_of(n / every), k = n % every;
      adjoint_step(&recs[b][2 * k], screcs[b] + 3 * k, torque, acc);
    } else {
      adjoint_step(nullptr, nullptr, torque, acc);
    }
  }
  end_record();
  const double T = get_scalars(0).t - t0;
  double I = 0.0;
  hip_check(hipMemcpy(&I, acc, 8, hipMemcpyDeviceToHost), "torque D2H");
  if (check_error() || !(T > 0.0)) return -1;
  *loss = I / T;
  const double wl = 1.0 / T, lamT = -I / (T * T);
  if (nseg > 1) {  // the state after nsteps, put back at the end
    snapshot(fin);
    launch_copy_scalars(fin_sc, d_sc, st);
  }

  // one substep's mesh: input state `in`, everything else from its record
  DevMesh base = mesh();
  base.Fface = aFS;
  auto sub_mesh = [&](const AdjSubRec &in, const AdjSubRec &o, DiskScalars *sc) {
    DevMesh m = base;
    m.prim = in.prim;
    m.tiph = in.tiph;
    m.dphi = in.dphi;
    m.prim_new = o.prim;
    m.tiph_new = o.tiph;
    m.dphi_new = o.dphi;
    m.grad = o.grad;
    m.gradp = o.gradp;
    m.fL = o.fL;
    m.fR = o.fR;
    m.fdphi = o.fdphi;
    m.fcm = o.fcm;
    m.fo_r = o.fo_r;
    m.run_start = o.run;
    m.pst = o.pst;
    m.u_save = o.usave;
    m.ftype = o.ftype;
    m.sc = sc;
    m.torque_tile = nullptr;
    return m;
  };

  // reverse sweep
  const DiskParams &P = hdp[0];
  AdjMesh a;
  a.aR = aR;
  a.au = au;
  a.agrad = agrad;
  a.agradp = agradp;
  a.aSL = aSL;
  a.aSR = aSR;
  a.fr = fr;
  a.part = part;
  a.npart = (long long)npart;
  a.wl = wl;
  a.lam_dt = &cs->lam_dt;
  a.visc = P.visc > 0.0;
  // the streams: the reverse sweep on st, the recomputes on st_aux.  A
  // recompute writes the engine's state and one record; it waits for the
  // loss pass and for the sweep that last read its record, and the sweep of
  // its segment waits for it.  The sweep reads only its record, its
  // checkpoint and the constant mesh (its face-slope scratch is aFS).
  const hipStream_t sm = st;
  hipEvent_t *const ev_rec = adj_ev, *const ev_back = adj_ev + 2, ev_fwd = adj_ev[4];
  bool used[2] = {true, true};  // the loss pass filled both records
  if (nbuf > 1) {
    hipEventRecord(ev_fwd, sm);
    hipStreamWaitEvent(st_aux, ev_fwd, 0);
  }
  // A segment's recompute from its checkpoint, recording it, in pieces: the
  // host issues one recompute step and one sweep step in turn, so that both
  // streams hold queued work (issued whole, the recompute's launches kept
  // the host busy for as long as the GPU took to run them, and the sweep
  // queued behind them found an idle stream only when they were done)
  auto recompute_begin = [&](int seg) {
    const int b = buf_of(seg);
    if (used[b]) hipStreamWaitEvent(st_aux, ev_back[b], 0);
    st = st_aux;
    adjoint_restore(ckpt + wA * seg, false);
    launch_copy_scalars(d_sc, ckpt_sc + seg, st);
    st = sm;
  };
  auto recompute_step = [&](int seg, int k) {  // step seg * every + k
    const int b = buf_of(seg);
    st = st_aux;
    adjoint_step(&recs[b][2 * k], screcs[b] + 3 * k, torque, acc);
    st = sm;
  };
  auto recompute_end = [&](int seg) {
    st = st_aux;
    end_record();
    st = sm;
    hipEventRecord(ev_rec[buf_of(seg)], st_aux);
  };
  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int n0 = seg * every, n1 = std::min(nsteps, n0 + every);
    const int b = buf_of(seg);
    // the previous segment's recompute (when the loss pass did not record
    // it) overlaps this segment's sweep; an earlier segment is whole
    const int rn = seg - 1 >= 0 && seg - 1 < nseg - nbuf ? every : 0;
    if (rn) recompute_begin(seg - 1);
    if (seg < nseg - nbuf) hipStreamWaitEvent(sm, ev_rec[b], 0);
    const std::vector<AdjSubRec> &rec = recs[b];
    DiskScalars *const screc = screcs[b];
    AdjSubRec start = rec[0];
    start.prim = ckpt + wA * seg;
    start.tiph = start.prim + nq;
    start.dphi = start.tiph + nz;
    for (int n = n1 - 1, r = 0; n >= n0 || r < rn; --n, ++r) {
      if (r < rn) recompute_step(seg - 1, r);
      if (n < n0) continue;
      const int k = n - n0;
      DiskScalars *scP = screc + 3 * k, *scA = scP + 1, *scB = scP + 2;
      const AdjSubRec &in1 = k == 0 ? start : rec[2 * k - 1];
      // substep 2 (fac 1/2)
      const DevMesh m2 = sub_mesh(rec[2 * k], rec[2 * k + 1], scB);
      a.aP = aP;
      a.aPn = aPn;
      a.fac = 0.5;
      a.copy_first = 0;
      a.part = part;
      a.arot = arot2;
      // riding in this substep's update launch: the rotation sums of the next
      // step's first substep (th1, used by this step's chain) and the step's
      // CFL cell (the step-start state with the scalars and tile minima the
      // forward's CFL minimum saw)
      AdjExtra x2;
      x2.rot_src = arot1;
      x2.pick_prim = in1.prim;
      x2.pick_tiph = in1.tiph;
      x2.pick_dphi = in1.dphi;
      x2.pick_dt_tile = rec[2 * k].dt_tile;
      x2.pick_sc = scP;
      x2.dtbits = &scA->dtmin_bits;
      x2.pick = pick;
      x2.rot_out = th1;
      launch_adjoint_substep(m2, a, plan, sm, x2);
      std::swap(aP, aPn);
      // substep 1 (fac 1, RKcons = prim2cons(prim))
      const DevMesh m1 = sub_mesh(in1, rec[2 * k], scA);
      a.aP = aP;
      a.aPn = aPn;
      a.fac = 1.0;
      a.copy_first = 1;
      a.part = part + 9 * npart;
      a.arot = arot1;
      // the second substep's rotation sums ride in this substep's update
      // launch, both substeps' scalar partials' reduction in its last
      AdjExtra x1;
      x1.rot_src = arot2;
      x1.rot_out = th2;
      x1.red_src = part;
      x1.red = red;
      launch_adjoint_substep(m1, a, plan, sm, x1);
      std::swap(aP, aPn);
      a.part = part;
      // the step's dt: its CFL minimum, at the step-start state with the
      // scalars the forward's CFL minimum saw
      DevMesh md = m1;
      md.sc = scP;
      a.aP = aPn;  // the step-start adjoint
      launch_adjoint_chain(md, a, pick, nann, scP, scA, scB, red, th1, th2, Ltheta, lamT, cs, sm);
    }
    if (rn) recompute_end(seg - 1);
    if (nbuf > 1) {
      hipEventRecord(ev_back[b], sm);
      used[b] = true;
    }
  }
  AdjChain C;
  hip_check(hipMemcpy(&C, cs, sizeof(C), hipMemcpyDeviceToHost), "adjoint D2H");
  hip_check(hipMemcpy(g_prim0, aPn, nq * 8, hipMemcpyDeviceToHost), "adjoint D2H");
  // (the blocking copies have waited for both streams; segment 0's
  // recompute, the last, precedes segment 0's sweep on st)
  if (nseg > 1) {
    adjoint_restore(fin, false);
    launch_copy_scalars(d_sc, fin_sc, st);
  }
  *g_alpha = C.Lalpha;
  // M0 = 1/(1+q), M1 = q/(1+q), r0 = M1, r1 = M0 (GravMass_init.c:53-84; the
  // binary keeps them), so each derivative is +-1/(1+q)^2
  double gq = 0.0;
  if (P.grav_type == 2) {
    const double q = sims[0]->MassRatio;
    gq = (C.Lr[0] - C.Lr[1] - C.LM[0] + C.LM[1]) / ((1. + q) * (1. + q));
  }
  *g_q = gq;
  return check_error() ? -1 : 0;
}
