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

DevMesh Engine::mesh_prev() const {
  DevMesh m = mesh();
  m.prim = d_prim[1 - cur];
  m.prim_new = d_prim[cur];
  m.tiph = d_tiph[1 - cur];
  m.tiph_new = d_tiph[cur];
  m.dphi = d_dphi[1 - cur];
  m.dphi_new = d_dphi[cur];
  return m;
}

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
  if (why) {
    disco_set_error(12, std::string("the adjoint supports ") + why);
    return false;
  }
  return true;
}

// One RK2 step exactly as enqueue_rk2 launches it, with the pre-substep
// scalar ops as their own launch so the state can be recorded between them.
// snapA / snapB (when non-null) receive the state at the start of each
// substep and rec its scalars; every k_update adds its torque to acc.
void Engine::adjoint_step(AdjStepRec *rec, double *snapA, double *snapB, double *torque, double *acc) {
  const size_t nz = (size_t)Ntot, nq = (size_t)NQ * Ntot;
  auto snap = [&](double *dst, bool rk) {
    hipMemcpyAsync(dst, d_prim[cur], nq * 8, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(dst + nq, d_tiph[cur], nz * 8, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(dst + nq + nz, d_dphi[cur], nz * 8, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(dst + nq + 2 * nz, d_RKtiph, nz * 8, hipMemcpyDeviceToDevice, st);
    if (rk) hipMemcpyAsync(dst + nq + 3 * nz, d_RKcons, nq * 8, hipMemcpyDeviceToDevice, st);
  };
  DevMesh m = mesh();
  launch_set_w(m, plan, st);
  launch_mindt(m, plan, st, NQ);
  if (rec) rec->pre = get_scalars(0);
  launch_scalar(m, plan, st, SC_STEP_BEGIN | SC_SUB_BEGIN);
  if (rec) {
    snap(snapA, false);
    rec->A = get_scalars(0);
  }
  adj_torque = torque;
  enqueue_substep(1.0, 1, 0, 0, SC_AFTER_UPDATE | SC_SUB1_END | SC_SUB_BEGIN);
  launch_torque_sum(torque, plan.n_utiles, acc, st);
  if (rec) {
    snap(snapB, true);
    rec->B = get_scalars(0);
  }
  enqueue_substep(0.5, 0, 1, 0, SC_AFTER_UPDATE | SC_SUB2_END);
  launch_torque_sum(torque, plan.n_utiles, acc, st);
  adj_torque = nullptr;
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
  for (int n = 0; n < nsteps; ++n) adjoint_step(nullptr, nullptr, nullptr, torque, acc);
  const double T = get_scalars(0).t - t0;
  double I = 0.0;
  hip_check(hipMemcpy(&I, acc, 8, hipMemcpyDeviceToHost), "torque D2H");
  hipFree(torque);
  hipFree(acc);
  *loss = T > 0.0 ? I / T : 0.0;
  return check_error() ? -1 : 0;
}

int Engine::adjoint(int nsteps, int every, double *loss, double *g_alpha, double *g_q, double *g_prim0) {
  if (!ok || !adjoint_scope()) return -1;
  if (nsteps < 1) {
    disco_set_error(12, "adjoint: nsteps must be >= 1");
    return -1;
  }
  if (every < 1 || every > nsteps) every = nsteps;
  const size_t nz = (size_t)Ntot, nq = (size_t)NQ * Ntot, nf = (size_t)NQ * NFtot;
  const size_t wA = nq + 3 * nz, wB = 2 * nq + 3 * nz;  // doubles per snapshot
  const int nseg = (nsteps + every - 1) / every;
  const int nz_t = plan.n_ztiles, nu_t = plan.n_utiles, nrf_t = plan.n_rftiles, ni_t = plan.n_itiles;
  const size_t npart = std::max((size_t)(nz_t + 2 * nu_t + 2 * nrf_t), (size_t)ni_t);

  double *aP, *aPn, *aR, *au, *agrad, *agradp, *aSL, *aSR, *fr, *usave, *torque, *acc, *ckpt, *snapA, *snapB;
  double *arot, *rotout;
  signed char *ftype;
  int *pick;
  ScalarAdj *part, *out;
  DiskScalars *sc_tmp;
  bool good = dalloc0(&aP, nq, "adjoint") && dalloc0(&aPn, nq, "adjoint") && dalloc0(&aR, nq, "adjoint") &&
              dalloc0(&au, nq, "adjoint") && dalloc0(&agrad, nq, "adjoint") && dalloc0(&agradp, nq, "adjoint") &&
              dalloc0(&aSL, nf, "adjoint") && dalloc0(&aSR, nf, "adjoint") &&
              dalloc0(&fr, (size_t)AFR_N * NFtot, "adjoint") && dalloc0(&usave, nq, "adjoint") &&
              dalloc0(&torque, (size_t)plan.n_utiles, "adjoint") && dalloc0(&acc, 1, "adjoint") &&
              dalloc0(&ckpt, wA * nseg, "adjoint checkpoints") && dalloc0(&snapA, wA * every, "adjoint steps") &&
              dalloc0(&snapB, wB * every, "adjoint steps") && dalloc0(&part, npart, "adjoint") &&
              dalloc0(&out, 1, "adjoint") && dalloc0(&sc_tmp, 1, "adjoint") && dalloc0(&arot, nz, "adjoint") &&
              dalloc0(&rotout, (size_t)nann, "adjoint") && dalloc0(&ftype, (size_t)NFtot, "adjoint") && dalloc0(&pick, 1, "adjoint");
  if (!good) return -1;
  std::vector<DiskScalars> ckpt_sc(nseg);
  std::vector<AdjStepRec> rec(every);

  // forward: the loss, with a checkpoint at the start of every segment
  const double t0 = get_scalars(0).t;
  for (int n = 0; n < nsteps; ++n) {
    if (n % every == 0) {
      double *c = ckpt + wA * (n / every);
      hipMemcpyAsync(c, d_prim[cur], nq * 8, hipMemcpyDeviceToDevice, st);
      hipMemcpyAsync(c + nq, d_tiph[cur], nz * 8, hipMemcpyDeviceToDevice, st);
      hipMemcpyAsync(c + nq + nz, d_dphi[cur], nz * 8, hipMemcpyDeviceToDevice, st);
      hipMemcpyAsync(c + nq + 2 * nz, d_RKtiph, nz * 8, hipMemcpyDeviceToDevice, st);
      ckpt_sc[n / every] = get_scalars(0);
    }
    adjoint_step(nullptr, nullptr, nullptr, torque, acc);
  }
  const double T = get_scalars(0).t - t0;
  double I = 0.0;
  hip_check(hipMemcpy(&I, acc, 8, hipMemcpyDeviceToHost), "torque D2H");
  if (check_error() || !(T > 0.0)) return -1;
  *loss = I / T;
  const double wl = 1.0 / T, lamT = -I / (T * T);
  // each annulus turns by Omega = (wiph + rOm_a) / r per unit time
  // (cell_update_phi, cell_misc.c:73-75; rOm_a = 0 here); wiph is analytic
  // for the mesh motions the adjoint accepts
  std::vector<double> wi(nann), Omega(nann), Ltheta(nann, 0.0), th1(nann), th2(nann);
  hip_check(hipMemcpy(wi.data(), d_wiph, nann * 8, hipMemcpyDeviceToHost), "wiph D2H");
  for (int a2 = 0; a2 < nann; ++a2) Omega[a2] = wi[a2] / (.5 * (hann[a2].rp + hann[a2].rm));

  // reverse sweep
  const DiskParams &P = hdp[0];
  double Lphi[2] = {0, 0}, Lr[2] = {0, 0}, LM[2] = {0, 0}, Lalpha = 0.0;
  ScalarAdj S0, S1, S2;
  auto fetch = [&](ScalarAdj &s) { hip_check(hipMemcpy(&s, out, sizeof(s), hipMemcpyDeviceToHost), "adjoint D2H"); };
  // adjoints of the mass phases and radii through xbh = r cos(phi), ybh = r sin(phi)
  auto chain_pos = [&](const DiskScalars &sc, const ScalarAdj &s, double *lphi, double *lr) {
    for (int p = 0; p < 2; ++p) {
      lphi[p] = -sc.ybh[p] * s.xbh[p] + sc.xbh[p] * s.ybh[p];
      const double r = sc.gm[p].r;
      lr[p] = r != 0.0 ? (sc.xbh[p] * s.xbh[p] + sc.ybh[p] * s.ybh[p]) / r : 0.0;
    }
  };
  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int n0 = seg * every, n1 = std::min(nsteps, n0 + every);
    adjoint_restore(ckpt + wA * seg, false);
    put_scalars(0, ckpt_sc[seg]);
    for (int n = n0; n < n1; ++n)
      adjoint_step(&rec[n - n0], snapA + wA * (n - n0), snapB + wB * (n - n0), torque, acc);
    for (int n = n1 - 1; n >= n0; --n) {
      const AdjStepRec &R = rec[n - n0];
      AdjMesh a;
      a.aR = aR;
      a.au = au;
      a.agrad = agrad;
      a.agradp = agradp;
      a.aSL = aSL;
      a.aSR = aSR;
      a.fr = fr;
      a.part = part;
      a.wl = wl;
      a.lam_dt = 0.0;
      a.arot = arot;
      a.visc = P.visc > 0.0;
      // substep 2 (fac 1/2, RK 1/2)
      adjoint_restore(snapB + wB * (n - n0), true);
      put_scalars(0, R.B);
      adj_usave = usave;  // k_update records what cons2prim converts
      adj_ftype = ftype;  // the face builder records the face types
      enqueue_substep(0.5, 0, 1, 0, 0);
      adj_usave = nullptr;
      adj_ftype = nullptr;
      DevMesh mb = mesh_prev();
      mb.u_save = usave;
      mb.ftype = ftype;
      hipMemsetAsync(arot, 0, nz * 8, st);
      a.aP = aP;
      a.aPn = aPn;
      a.fac = 0.5;
      a.RK = R.B.RK;
      a.copy_first = 0;
      launch_adjoint_substep(mb, a, plan, st, out);
      launch_adjoint_rot(mb, a, plan, st, rotout);
      fetch(S2);
      hip_check(hipMemcpy(th2.data(), rotout, nann * 8, hipMemcpyDeviceToHost), "adjoint D2H");
      std::swap(aP, aPn);
      // substep 1 (fac 1, RK 0, RKcons = prim2cons(prim))
      adjoint_restore(snapA + wA * (n - n0), false);
      put_scalars(0, R.A);
      adj_usave = usave;
      adj_ftype = ftype;
      enqueue_substep(1.0, 1, 0, 0, 0);
      adj_usave = nullptr;
      adj_ftype = nullptr;
      mb = mesh_prev();
      mb.u_save = usave;
      mb.ftype = ftype;
      hipMemsetAsync(arot, 0, nz * 8, st);
      a.aP = aP;
      a.aPn = aPn;
      a.fac = 1.0;
      a.RK = R.A.RK;
      a.copy_first = 1;
      launch_adjoint_substep(mb, a, plan, st, out);
      fetch(S1);
      std::swap(aP, aPn);
      // dt of the step: both substeps' dt_s, the phase it advances, T
      const double omega = R.A.gm[0].omega;
      double lphi1[2], lr1[2], lphi2[2], lr2[2], lphi0[2] = {0, 0}, lr0[2] = {0, 0};
      chain_pos(R.A, S1, lphi1, lr1);
      chain_pos(R.B, S2, lphi2, lr2);
      double lam_dt = 1.0 * S1.dt + 0.5 * S2.dt + omega * (Lphi[0] + Lphi[1] + lphi2[0] + lphi2[1]) + lamT;
      for (int a2 = 0; a2 < nann; ++a2) lam_dt += Omega[a2] * (Ltheta[a2] + th2[a2]);
      memset(&S0, 0, sizeof(S0));
      double dtmin;
      memcpy(&dtmin, &R.A.dtmin_bits, 8);
      if (R.A.dt == dtmin) {  // not clipped at T_MAX
        hip_check(hipMemcpyAsync(sc_tmp, &R.pre, sizeof(DiskScalars), hipMemcpyHostToDevice, st), "adjoint");
        DevMesh md = mesh_prev();  // step-start state, scalars as k_mindt saw them
        md.sc = sc_tmp;
        md.ftype = ftype;
        a.aP = aPn;  // the step-start adjoint
        a.lam_dt = lam_dt;
        launch_adjoint_dt(md, a, plan, st, R.A.dtmin_bits, pick, out);
        fetch(S0);
        chain_pos(R.pre, S0, lphi0, lr0);
      }
      launch_adjoint_rot(mb, a, plan, st, rotout);  // substep 1 and the CFL step
      hip_check(hipMemcpy(th1.data(), rotout, nann * 8, hipMemcpyDeviceToHost), "adjoint D2H");
      for (int a2 = 0; a2 < nann; ++a2) Ltheta[a2] += th1[a2] + th2[a2];
      for (int p = 0; p < 2; ++p) {
        Lphi[p] += lphi1[p] + lphi2[p] + lphi0[p];
        Lr[p] += lr1[p] + lr2[p] + lr0[p];
        LM[p] += S1.M[p] + S2.M[p] + S0.M[p];
      }
      Lalpha += S1.alpha + S2.alpha + S0.alpha;
    }
  }
  hip_check(hipMemcpy(g_prim0, aPn, nq * 8, hipMemcpyDeviceToHost), "adjoint D2H");
  *g_alpha = Lalpha;
  // M0 = 1/(1+q), M1 = q/(1+q), r0 = M1, r1 = M0 (GravMass_init.c:53-84; the
  // binary keeps them), so each derivative is +-1/(1+q)^2
  double gq = 0.0;
  if (P.grav_type == 2) {
    const double q = sims[0]->MassRatio;
    gq = (Lr[0] - Lr[1] - LM[0] + LM[1]) / ((1. + q) * (1. + q));
  }
  *g_q = gq;
  for (void *p : {(void *)aP, (void *)aPn, (void *)aR, (void *)au, (void *)agrad, (void *)agradp, (void *)aSL,
                  (void *)aSR, (void *)fr, (void *)usave, (void *)torque, (void *)acc, (void *)ckpt, (void *)snapA,
                  (void *)snapB, (void *)part, (void *)out, (void *)sc_tmp, (void *)arot, (void *)rotout,
                  (void *)ftype, (void *)pick})
    hipFree(p);
  return check_error() ? -1 : 0;
}
