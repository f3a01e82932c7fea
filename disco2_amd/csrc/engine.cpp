// Engine construction, host<->device transfers and the step sequence.
#include "engine.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <tuple>
#include <string>

#include "../../include/disco.h"
#include "initial_data.h"

bool hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  disco_set_error(20, std::string(what) + ": " + hipGetErrorString(e));
  return false;
}

template <typename T>
static bool dalloc(T **p, size_t n, const char *what) {
  if (n == 0) n = 1;
  return hip_check(hipMalloc((void **)p, n * sizeof(T)), what);
}
template <typename T>
static bool h2d(T *d, const T *h, size_t n) {
  if (n == 0) return true;
  return hip_check(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
}

Engine::Engine(int n, Sim **s) : ndisk(n), sims(s, s + n) {
  NQ = sims[0]->NUM_Q;
  for (int d = 0; d < n; ++d) {
    if (sims[d]->NUM_Q != NQ) {
      disco_set_error(7, "all disks of a batch need the same NUM_Q");
      return;
    }
  }
  if (NQ < 5 || NQ > DISCO_MAXQ) {
    disco_set_error(7, "NUM_Q must be 5 to 8 (NUM_C 5 with at most three passive scalars)");
    return;
  }
  std::vector<double> hrf, hzf;
  std::vector<int2> ztiles, utiles, stiles, itiles, rft, zft, bc0, bc1, szt, bit;
  std::vector<InitRow> hinit;
  bool any_init = false;
  long long zb = 0, fb = 0;
  int rb = 0;
  // r pairs of all disks first, then z pairs
  std::vector<PairInfo> zpairs;
  for (int d = 0; d < n; ++d) {
    Sim *S = sims[d];
    if (S->NUM_C != 5) {
      disco_set_error(7, "NUM_C must be 5");
      return;
    }
    if (S->W_NUMERIC_TYPE == C_WCELL) {
      disco_set_error(8, "w_numeric_type 1 (C_WCELL) is rejected by the reference too");
      return;
    }
    DiskParams P;
    memset(&P, 0, sizeof(P));
    sim_param_arrays(S, &P);
    int NR = P.NR, NZ = P.NZ;
    if (NR < 4) {
      disco_set_error(8, "need at least one interior annulus");
      return;
    }
    for (int i = 0; i < NR; ++i)
      if (S->N_p[i] < 2) {
        disco_set_error(8, "every annulus needs N_p >= 2 (check that R_Min exceeds the radial spacing)");
        return;
      }
    P.geo_r = (int)hrf.size();
    P.geo_z = (int)hzf.size();
    hrf.insert(hrf.end(), S->r_faces.begin(), S->r_faces.end());
    hzf.insert(hzf.end(), S->z_faces.begin(), S->z_faces.end());
    long long NL = 0;
    std::vector<long long> aoff(NR + 1, 0);
    for (int i = 0; i < NR; ++i) aoff[i + 1] = aoff[i] + S->N_p[i];
    NL = aoff[NR];
    P.NL = (int)NL;
    P.zone_base = zb;
    P.ann_base = (int)hann.size();
    P.itile0 = (int)itiles.size();
    P.utile0 = (int)utiles.size();
    zbase.push_back(zb);
    nzones.push_back(NL * NZ);
    // fixed r boundaries and damping: rows of the single-cell initial data
    std::vector<InitRow> rows;
    bool any = false;
    if (!disco_init_rows(S, P.NQ, rows, &any)) return;
    hinit.insert(hinit.end(), rows.begin(), rows.end());
    any_init |= any;
    for (int k = 0; k < NZ; ++k)
      for (int i = 0; i < NR; ++i) {
        AnnInfo A;
        memset(&A, 0, sizeof(A));
        A.zbase = zb + k * NL + aoff[i];
        A.N = S->N_p[i];
        A.disk = d;
        A.k = k;
        A.i = i;
        A.rm = S->r_faces[i];
        A.rp = S->r_faces[i + 1];
        A.zm = S->z_faces[k];
        A.zp = S->z_faces[k + 1];
        bool interior = i >= P.NgMinR && i < NR - P.NgMaxR && k >= P.NgMinZ && k < NZ - P.NgMaxZ;
        A.itile = interior ? (int)itiles.size() : -1;
        A.utile = (int)utiles.size();
        int a = (int)hann.size();
        hann.push_back(A);
        for (int j0 = 0; j0 < A.N; j0 += DISCO_TILE) ztiles.push_back(make_int2(a, j0));
        for (int j0 = 0; j0 < A.N; j0 += DISCO_TILE - 1) utiles.push_back(make_int2(a, j0));
        for (int j0 = 0; j0 < A.N; j0 += DISCO_TILE - 2) stiles.push_back(make_int2(a, j0));
        if (interior)
          for (int j0 = 0; j0 < A.N; j0 += DISCO_TILE) itiles.push_back(make_int2(a, j0));
        if (P.bound_r == BOUND_OUTFLOW) {
          if ((i == 0 && P.no_inner_bc != 1) || i == NR - 2)
            for (int j0 = 0; j0 < A.N; j0 += DISCO_TILE) bc0.push_back(make_int2(a, j0));
          if (i == NR - 1)
            for (int j0 = 0; j0 < A.N; j0 += DISCO_TILE) bc1.push_back(make_int2(a, j0));
        }
        if (P.NglobalZ > 1 && (k < P.NgMinZ || k >= NZ - P.NgMaxZ))
          for (int j0 = 0; j0 < A.N; j0 += DISCO_TILE) szt.push_back(make_int2(a, j0));
        if (rows[(size_t)k * NR + i].flags & (INIT_FIXED | INIT_DAMP))
          for (int j0 = 0; j0 < A.N; j0 += DISCO_TILE) bit.push_back(make_int2(a, j0));
      }
    P.itile1 = (int)itiles.size();
    P.utile1 = (int)utiles.size();
    zb += NL * NZ;
    // r pairs (k, i)-(k, i+1): constants of build_jloop, src/Face/face_create_destroy.c:60-68
    P.rpair_base = (int)hpairs.size();
    for (int k = 0; k < NZ; ++k)
      for (int i = 0; i < NR - 1; ++i) {
        PairInfo Q;
        memset(&Q, 0, sizeof(Q));
        Q.ann_in = P.ann_base + k * NR + i;
        Q.ann_out = P.ann_base + k * NR + i + 1;
        Q.disk = d;
        Q.dir = 0;
        int Nin = S->N_p[i], Nout = S->N_p[i + 1];
        Q.cap = Nin + 2 * Nout + 4;
        Q.run_cap = 2 * Nout + 4;
        Q.deltaL = .5 * (sim_FacePos(S, i, 0) - sim_FacePos(S, i - 1, 0));
        Q.deltaR = .5 * (sim_FacePos(S, i + 1, 0) - sim_FacePos(S, i, 0));
        Q.r = sim_FacePos(S, i, 0);
        double zp = sim_FacePos(S, k, 1), zm = sim_FacePos(S, k - 1, 1);
        Q.deltaPerp = zp - zm;
        hpairs.push_back(Q);
      }
    // z pairs (k, i)-(k+1, i): src/Face/face_create_destroy.c:70-78
    if (P.NglobalZ > 1) {
      P.zpair_base = (int)zpairs.size();  // rebased below
      for (int k = 0; k < NZ - 1; ++k)
        for (int i = 0; i < NR; ++i) {
          PairInfo Q;
          memset(&Q, 0, sizeof(Q));
          Q.ann_in = P.ann_base + k * NR + i;
          Q.ann_out = P.ann_base + (k + 1) * NR + i;
          Q.disk = d;
          Q.dir = 1;
          int N = S->N_p[i];
          Q.cap = 3 * N + 4;
          Q.run_cap = 2 * N + 4;
          Q.deltaL = .5 * (sim_FacePos(S, k, 1) - sim_FacePos(S, k - 1, 1));
          Q.deltaR = .5 * (sim_FacePos(S, k + 1, 1) - sim_FacePos(S, k, 1));
          double rp = sim_FacePos(S, i, 0), rm = sim_FacePos(S, i - 1, 0);
          Q.deltaPerp = rp - rm;
          Q.r = .5 * (rp + rm);
          zpairs.push_back(Q);
        }
    }
    hdp.push_back(P);
  }
  n_rpairs = (int)hpairs.size();
  n_zpairs = (int)zpairs.size();
  for (auto &P : hdp)
    if (P.NglobalZ > 1) P.zpair_base += n_rpairs;
  hpairs.insert(hpairs.end(), zpairs.begin(), zpairs.end());
  for (int p = 0; p < (int)hpairs.size(); ++p) {
    PairInfo &Q = hpairs[p];
    Q.face_base = fb;
    Q.run_base = rb;
    fb += Q.cap;
    rb += Q.run_cap;
    for (int f0 = 0; f0 < Q.cap; f0 += DISCO_TILE) (Q.dir == 0 ? rft : zft).push_back(make_int2(p, f0));
  }
  Ntot = zb;
  NFtot = fb;
  w_static = true;
  for (auto &P : hdp) w_static = w_static && P.w_numeric != C_RIGID;
  nrun = rb > 0 ? rb : 1;
  nann = (int)hann.size();

  bool good = true;
  good &= hip_check(hipStreamCreate(&st), "hipStreamCreate");
  good &= hip_check(hipEventCreate(&ev0), "hipEventCreate") && hip_check(hipEventCreate(&ev1), "hipEventCreate");
  good &= dalloc(&d_dp, hdp.size(), "dp") && h2d(d_dp, hdp.data(), hdp.size());
  good &= dalloc(&d_sc, (size_t)n, "scalars");
  good &= dalloc(&d_ann, hann.size(), "ann") && h2d(d_ann, hann.data(), hann.size());
  good &= dalloc(&d_pairs, hpairs.size(), "pairs") && h2d(d_pairs, hpairs.data(), hpairs.size());
  good &= dalloc(&d_pst, hpairs.size(), "pair state");
  good &= dalloc(&d_rf, hrf.size(), "rf") && h2d(d_rf, hrf.data(), hrf.size());
  good &= dalloc(&d_zf, hzf.size(), "zf") && h2d(d_zf, hzf.data(), hzf.size());
  size_t nq = (size_t)NQ * Ntot;
  good &= dalloc(&d_prim[0], nq, "prim") && dalloc(&d_prim[1], nq, "prim") && dalloc(&d_RKcons, nq, "RKcons");
  good &= dalloc(&d_grad, nq, "grad") && dalloc(&d_gradp, nq, "gradp");
  if (n_zpairs) good &= dalloc(&d_grad_r, nq, "grad_r");  // both slope directions in one pass
  good &= dalloc(&d_tiph[0], Ntot, "tiph") && dalloc(&d_tiph[1], Ntot, "tiph");
  good &= dalloc(&d_RKtiph, Ntot, "RKtiph") && dalloc(&d_dphi[0], Ntot, "dphi") && dalloc(&d_dphi[1], Ntot, "dphi");
  good &= dalloc(&d_wiph, (size_t)nann, "wiph");
  good &= dalloc(&d_fL, (size_t)NFtot, "fL") && dalloc(&d_fR, (size_t)NFtot, "fR");
  good &= dalloc(&d_fdphi, (size_t)NFtot, "fdphi") && dalloc(&d_fcm, (size_t)NFtot, "fcm");
  good &= dalloc(&d_fo_r, Ntot, "fo_r") && dalloc(&d_fo_z, Ntot, "fo_z");
  good &= dalloc(&d_run, (size_t)(rb > 0 ? rb : 1), "run_start");
  good &= dalloc(&d_Fface, (size_t)(NQ + 3) * NFtot, "Fface");
  good &= dalloc(&d_err, 1, "err");
  good &= dalloc(&d_dt_tile, itiles.empty() ? 1 : itiles.size(), "dt tiles");
  good &= dalloc(&d_ann_rot, (size_t)nann, "annulus rotation");
  bool any_sink = false;
  for (auto &P : hdp) any_sink |= P.sink == 1;
  if (any_sink) good &= dalloc(&d_mdot_tile, 2 * utiles.size(), "sink tiles");
  if (any_init) good &= dalloc(&d_init, hinit.size(), "init rows") && h2d(d_init, hinit.data(), hinit.size());
  if (!good) return;
  hipMemset(d_err, 0, sizeof(int));
  hipMemset(d_prim[0], 0, nq * sizeof(double));
  hipMemset(d_prim[1], 0, nq * sizeof(double));
  hipMemset(d_RKcons, 0, nq * sizeof(double));
  hipMemset(d_grad, 0, nq * sizeof(double));
  hipMemset(d_gradp, 0, nq * sizeof(double));
  hipMemset(d_wiph, 0, (size_t)nann * sizeof(double));
  hipMemset(d_ann_rot, 0, (size_t)nann * sizeof(double4));
  {  // no faces yet: every gather over a pair is empty until k_faces runs
    std::vector<PairState> p0(hpairs.size());
    for (auto &x : p0) x = PairState{0, 0, -1, 0, 0};
    h2d(d_pst, p0.data(), p0.size());
  }
  hipMemset(d_fo_r, 0, Ntot * sizeof(int));
  hipMemset(d_fo_z, 0, Ntot * sizeof(int));
  hipMemset(d_run, 0, (size_t)(rb > 0 ? rb : 1) * sizeof(int));
  std::vector<DiskScalars> sc0(n);
  memset(sc0.data(), 0, sizeof(DiskScalars) * n);
  for (auto &s : sc0) {
    double big = 1.e100;
    memcpy(&s.dtmin_bits, &big, 8);
  }
  h2d(d_sc, sc0.data(), (size_t)n);
  {
    std::vector<double> none(itiles.empty() ? 1 : itiles.size(), 1.e100);
    h2d(d_dt_tile, none.data(), none.size());
  }

  // the table-order lists the layer-block schedules reorder (build_zblock)
  h_ztiles = ztiles;
  h_utiles = utiles;
  h_zft = zft;
  // one tile array on the device, sliced per list
  std::vector<int2> all;
  auto add = [&](std::vector<int2> &v, int &count) {
    count = (int)v.size();
    size_t off = all.size();
    all.insert(all.end(), v.begin(), v.end());
    return off;
  };
  size_t o_z = add(ztiles, plan.n_ztiles), o_u = add(utiles, plan.n_utiles), o_i = add(itiles, plan.n_itiles), o_r = add(rft, plan.n_rftiles),
         o_zf = add(zft, plan.n_zftiles), o_b0 = add(bc0, plan.n_bc0), o_b1 = add(bc1, plan.n_bc1),
         o_s = add(szt, plan.n_sz), o_bi = add(bit, plan.n_binit), o_st = add(stiles, plan.n_stiles);
  if (!dalloc(&d_tiles, all.size(), "tiles") || !h2d(d_tiles, all.data(), all.size())) return;
  plan.ztiles = d_tiles + o_z;
  plan.utiles = d_tiles + o_u;
  plan.itiles = d_tiles + o_i;
  plan.rftiles = d_tiles + o_r;
  plan.zftiles = d_tiles + o_zf;
  plan.bc0 = d_tiles + o_b0;
  plan.bc1 = d_tiles + o_b1;
  plan.sz = d_tiles + o_s;
  plan.binit = d_tiles + o_bi;
  plan.stiles = d_tiles + o_st;
  tl_zb[0][0] = plan.ztiles;
  tl_zb[0][1] = plan.utiles;
  tl_zb[0][2] = plan.zftiles;
  plan.ndisk = n;
  plan.any_visc = 0;
  plan.feat = 0;
  for (auto &P : hdp) {
    plan.any_visc |= P.visc > 0.0;
    plan.feat |= (P.visc > 0.0 ? F_VISC : 0) | (P.riemann == HLL_SOLVER ? F_HLL : F_HLLC) |
                 (P.sink == 1 ? F_SINK : 0) | (P.cooling == 1 ? F_COOL : 0) |
                 (P.w_analytic != A_FIXED ? F_FRAME : 0);
  }
  plan.nann = nann;
  {  // largest outer annulus of each pair direction, staged in LDS if it fits
    int mr = 0, mz = 0;
    for (auto &Q : hpairs) {
      int N = hann[Q.ann_out].N;
      if (Q.dir == 0)
        mr = N > mr ? N : mr;
      else
        mz = N > mz ? N : mz;
    }
    const int cap = FACE_STAGE_BYTES / (2 * (int)sizeof(double));
    plan.stage_r = mr <= cap ? mr : 0;
    plan.stage_z = mz <= cap ? mz : 0;
    if (!prepare_launches(plan)) plan.stage_r = plan.stage_z = 0;
    stage_r_max = plan.stage_r;
    stage_z_max = plan.stage_z;
  }
  {  // the face builder's pairs, largest outer annulus first (stable), per direction
    std::vector<int> perm(hpairs.size());
    for (size_t q = 0; q < perm.size(); ++q) perm[q] = (int)q;
    auto bigger = [&](int x, int y) { return hann[hpairs[x].ann_out].N > hann[hpairs[y].ann_out].N; };
    std::stable_sort(perm.begin(), perm.begin() + n_rpairs, bigger);
    std::stable_sort(perm.begin() + n_rpairs, perm.end(), bigger);
    if (!dalloc(&d_pair_perm, perm.empty() ? 1 : perm.size(), "pair order") ||
        (!perm.empty() && !h2d(d_pair_perm, perm.data(), perm.size())))
      return;
  }
  plan.n_rpairs = n_rpairs;
  plan.n_zpairs = n_zpairs;
  // frames that do not depend on the binary separation: tabulated on the
  // host once (k_scalar recomputes the A_MILOS frame when the masses move)
  for (int d = 0; d < n; ++d) {
    const DiskParams &P = hdp[d];
    if (P.w_analytic == A_FIXED || P.w_analytic == A_MILOS) continue;
    static_assert(sizeof(double4) == 4 * sizeof(double), "double4 rows");
    std::vector<double> rows;
    disco_frame_rows(sims[d], 1.0, rows);
    if (!hip_check(hipMemcpy(d_ann_rot + P.ann_base, rows.data(), rows.size() * sizeof(double),
                             hipMemcpyHostToDevice), "frame rows"))
      return;
  }
  // initial mesh edges: cell_create + cell_clean_pi + tiph sync (src/main.c:77-85)
  for (int d = 0; d < n; ++d) {
    std::vector<double> t, dp;
    disco_cell_offsets(sims[d], t, dp);
    upload(d, 5, t.data());
    upload(d, 7, dp.data());
    upload(d, 6, t.data());
  }
  ok = hip_check(hipDeviceSynchronize(), "engine set-up");
}

Engine::~Engine() {
  if (gexec) hipGraphExecDestroy(gexec);
  if (graph) hipGraphDestroy(graph);
  void *ptrs[] = {d_dp, d_sc, d_ann, d_pairs, d_pst, d_rf, d_zf, d_prim[0], d_prim[1], d_RKcons, d_grad, d_gradp, d_grad_r,
                  d_tiph[0], d_tiph[1], d_RKtiph, d_dphi[0], d_dphi[1], d_wiph, d_fL, d_fR, d_fdphi, d_fcm, d_fo_r,
                  d_fo_z, d_run, d_Fface, d_err, d_tiles, d_tiles_zb[1], d_tiles_zb[2], d_tiles_zb[3],
                  d_tiles_zb[4], d_tiles_zb[5], d_tiles_zb[6], d_tiles_zb[7], d_dt_tile, d_mdot_tile, d_ann_rot, d_init, d_pair_perm, adj_mem, diag_mem,
                  ckpt_mem, d_ucons, d_cool, d_umdot, d_urows, d_led_torque};
  for (void *p : ptrs)
    if (p) hipFree(p);
  if (ev0) hipEventDestroy(ev0);
  if (ev1) hipEventDestroy(ev1);
  if (st_aux) hipStreamDestroy(st_aux);
  for (hipEvent_t e : adj_ev)
    if (e) hipEventDestroy(e);
  if (st) hipStreamDestroy(st);
}

DevMesh Engine::mesh() const {
  DevMesh m;
  m.dp = d_dp;
  m.sc = d_sc;
  m.ann = d_ann;
  m.pairs = d_pairs;
  m.pst = d_pst;
  m.rf = d_rf;
  m.zf = d_zf;
  m.prim = d_prim[cur];
  m.prim_new = d_prim[1 - cur];
  m.RKcons = d_RKcons;
  m.grad = d_grad;
  m.grad_r = d_grad_r;
  m.gradp = d_gradp;
  m.tiph = d_tiph[cur];
  m.tiph_new = d_tiph[1 - cur];
  m.RKtiph = d_RKtiph;
  m.dphi = d_dphi[cur];
  m.dphi_new = d_dphi[1 - cur];
  m.wiph = d_wiph;
  m.fL = d_fL;
  m.fR = d_fR;
  m.fdphi = d_fdphi;
  m.fcm = d_fcm;
  m.fo_r = d_fo_r;
  m.fo_z = d_fo_z;
  m.run_start = d_run;
  m.Fface = d_Fface;
  m.Ntot = Ntot;
  m.NFtot = NFtot;
  m.err = d_err;
  m.dt_tile = d_dt_tile;
  m.mdot_tile = d_mdot_tile;
  m.ann_rot = d_ann_rot;
  m.init = d_init;
  m.sched_xcd = sched.xcd;
  m.pair_perm = sched.largest_first ? d_pair_perm : nullptr;
  m.u_save = adj_usave;
  m.torque_tile = adj_torque ? adj_torque : d_led_torque;
  m.torque_acc = adj_torque ? adj_torque_acc : nullptr;
  m.sc_rec = adj_sc_rec;
  m.ftype = adj_ftype;
  return m;
}

// field ids: 0 prim, 1 cons (download only), 2 RKcons, 3 grad, 4 gradp,
// 5 tiph, 6 RKtiph, 7 dphi, 8 wiph (download only, expanded per zone)
void Engine::upload(int d, int field, const double *src) {
  long long nz = nzones[d], z0 = zbase[d];
  auto put = [&](double *base, int nq) {
    for (int q = 0; q < nq; ++q)
      hip_check(hipMemcpy(base + q * Ntot + z0, src + q * nz, nz * sizeof(double), hipMemcpyHostToDevice),
                "upload");
  };
  switch (field) {
    case 0: put(d_prim[cur], NQ); break;
    case 2: put(d_RKcons, NQ); break;
    case 3: put(d_grad, NQ); break;
    case 4: put(d_gradp, NQ); break;
    case 5: put(d_tiph[cur], 1); break;
    case 6: put(d_RKtiph, 1); break;
    case 7: put(d_dphi[cur], 1); break;
    default: disco_set_error(9, "field cannot be uploaded");
  }
}

void Engine::download(int d, int field, double *dst) {
  long long nz = nzones[d], z0 = zbase[d];
  hip_check(hipStreamSynchronize(st), "synchronize");
  auto get = [&](const double *base, int nq, double *out) {
    for (int q = 0; q < nq; ++q)
      hip_check(hipMemcpy(out + q * nz, base + q * Ntot + z0, nz * sizeof(double), hipMemcpyDeviceToHost),
                "download");
  };
  switch (field) {
    case 0: get(d_prim[cur], NQ, dst); break;
    case 2: get(d_RKcons, NQ, dst); break;
    case 3: get(d_grad, NQ, dst); break;
    case 4: get(d_gradp, NQ, dst); break;
    case 5: get(d_tiph[cur], 1, dst); break;
    case 6: get(d_RKtiph, 1, dst); break;
    case 7: get(d_dphi[cur], 1, dst); break;
    case 8: {
      std::vector<double> w(nann);
      hip_check(hipMemcpy(w.data(), d_wiph, nann * sizeof(double), hipMemcpyDeviceToHost), "download");
      long long o = 0;
      for (int a = hdp[d].ann_base; a < hdp[d].ann_base + hdp[d].NR * hdp[d].NZ; ++a)
        for (int j = 0; j < hann[a].N; ++j) dst[o++] = w[a];
      break;
    }
    case 1: {  // prim2cons with cell_calc_cons volumes (src/Cell/cell_conversion.c:35-55)
      if (ucons_live && d_ucons) {  // the unit-level conserved state
        get(d_ucons, NQ, dst);
        break;
      }
      std::vector<double> pr((size_t)NQ * nz), dp(nz);
      get(d_prim[cur], NQ, pr.data());
      get(d_dphi[cur], 1, dp.data());
      Sim *S = sims[d];
      const DiskParams &P = hdp[d];
      for (int a = P.ann_base; a < P.ann_base + P.NR * P.NZ; ++a) {
        const AnnInfo &A = hann[a];
        double rp = A.rp, rm = A.rm, dz = A.zp - A.zm, r = .5 * (rp + rm);
        for (int j = 0; j < A.N; ++j) {
          long long z = A.zbase - z0 + j;
          double p[8], c[8];
          for (int q = 0; q < NQ; ++q) p[q] = pr[q * nz + z];
          double dV = .5 * (rp * rp - rm * rm) * dp[z] * dz;
          cell_prim2cons(p, c, r, dV, S);
          for (int q = 0; q < NQ; ++q) dst[q * nz + z] = c[q];
        }
      }
      break;
    }
    default: disco_set_error(9, "unknown field");
  }
}

void Engine::put_scalars(int d, const DiskScalars &s) {
  hip_check(hipMemcpy(d_sc + d, &s, sizeof(s), hipMemcpyHostToDevice), "scalars H2D");
}
DiskScalars Engine::get_scalars(int d) {
  DiskScalars s;
  hip_check(hipStreamSynchronize(st), "synchronize");
  hip_check(hipMemcpy(&s, d_sc + d, sizeof(s), hipMemcpyDeviceToHost), "scalars D2H");
  return s;
}

// src/main.c:103-119: gravMass init, clean_pi, z sync of the initial
// primitives and edges; conserved variables are derived on the fly.
void Engine::init_state() {
  for (int d = 0; d < ndisk; ++d) {
    Sim *S = sims[d];
    GravMass g[2];
    memset(g, 0, sizeof(g));
    (*gravMass_init_ptr(S))(g, S);
    gravMass_clean_pi(g, S);
    DiskScalars s;
    memset(&s, 0, sizeof(s));
    for (int p = 0; p < 2; ++p) {
      s.gm[p].r = g[p].r;
      s.gm[p].phi = g[p].phi;
      s.gm[p].M = g[p].M;
      s.gm[p].omega = g[p].omega;
      s.gm[p].Mdot = g[p].Mdot;
      s.gm[p].Macc = g[p].Macc;
    }
    double big = 1.e100;
    memcpy(&s.dtmin_bits, &big, 8);
    put_scalars(d, s);
    launch_scalar(mesh(), plan, st, SC_MASS);
    const DiskParams &P = hdp[d];
    if (P.NglobalZ > 1) {  // src/Cell/cell_sync.c:156-205 on one rank
      long long nz = nzones[d], NL = P.NL;
      std::vector<double> pr((size_t)NQ * nz), t(nz), rk((size_t)NQ * nz);
      download(d, 0, pr.data());
      download(d, 5, t.data());
      download(d, 2, rk.data());
      int NZ = P.NZ, kNm2g = NZ - 2 * P.NgMaxZ, kNmg = NZ - P.NgMaxZ, kg = P.NgMinZ;
      auto cp = [&](int kd, int ks) {
        for (long long x = 0; x < NL; ++x) {
          for (int q = 0; q < NQ; ++q) {
            pr[q * nz + kd * NL + x] = pr[q * nz + ks * NL + x];
            rk[q * nz + kd * NL + x] = rk[q * nz + ks * NL + x];
          }
          t[kd * NL + x] = t[ks * NL + x];
        }
      };
      for (int dd = 0; dd < P.NgMaxZ; ++dd) cp(kNmg + dd, kg + dd);
      for (int dd = 0; dd < P.NgMinZ; ++dd) cp(dd, kNm2g + dd);
      upload(d, 0, pr.data());
      upload(d, 5, t.data());
      upload(d, 2, rk.data());
    }
  }
  hip_check(hipDeviceSynchronize(), "init");
}

void Engine::enqueue_substep(double fac, int copy_first, int recompute_w, int pre_ops, int post_ops) {
  launch_substep(mesh(), plan, st, NQ, fac, copy_first, recompute_w, pre_ops, post_ops);
  cur = 1 - cur;
}

// timestep_rk2, src/TimeStep/timestep_algorithm.c:11-41.  The per-disk
// scalar updates between substeps ride in one launch after each update; the
// CFL tile minima of the step (timestep_set_dt) ride on the first substep's
// phi slopes, which read the same state.
void Engine::enqueue_rk2() {
  enqueue_set_w();
  enqueue_substep(1.0, 1, 0, SC_PLM_DT | SC_STEP_BEGIN | SC_SUB_BEGIN,
                  SC_AFTER_UPDATE | SC_SUB1_END | SC_SUB_BEGIN);
  enqueue_substep(0.5, 0, w_static ? 0 : 1, 0, SC_AFTER_UPDATE | SC_SUB2_END);
}

// cell_clear_w + cell_set_w (cell_set_w.c:12-102).  Every mesh motion but
// C_RIGID is a function of r alone: once computed, wiph is what set_w would
// write again, bit for bit, so later steps skip the launch.
void Engine::enqueue_set_w() {
  if (w_static && w_ready) return;
  launch_set_w(mesh(), plan, st);
  w_ready = true;
}

// timestep_forward_euler, src/TimeStep/timestep_algorithm.c:43-57
void Engine::enqueue_euler() {
  enqueue_set_w();
  enqueue_substep(1.0, 1, 0, SC_PLM_DT | SC_STEP_BEGIN | SC_SUB_BEGIN, SC_AFTER_UPDATE | SC_EULER_END);
}

void Engine::run(int nsteps, bool use_graph) {
  if (!ok) return;
  if (use_graph && (gexec == nullptr || graph_cur != cur)) {
    if (gexec) {
      hipGraphExecDestroy(gexec);
      hipGraphDestroy(graph);
      gexec = nullptr;
    }
    int cur0 = cur;
    if (w_static) enqueue_set_w();  // a mesh velocity of r alone: set once, before the capture, not in the graph
    if (!hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "capture")) return;
    enqueue_rk2();
    if (!hip_check(hipStreamEndCapture(st, &graph), "capture end")) return;
    if (!hip_check(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0), "graph instantiate")) return;
    cur = cur0;  // capture does not execute; a step swaps the edge buffers twice
    graph_cur = cur;
  }
  if (use_graph && w_static) enqueue_set_w();  // the graph holds no set_w; a no-op once wiph is set
  hip_check(hipEventRecord(ev0, st), "event");
  for (int s = 0; s < nsteps; ++s) {
    if (use_graph)
      hip_check(hipGraphLaunch(gexec, st), "graph launch");
    else
      enqueue_rk2();
  }
  hip_check(hipEventRecord(ev1, st), "event");
}

// Layer-block order (schedule field zb): the zone, update and z-face tiles of
// each disk listed in blocks of 2^zb layers, annulus column by annulus
// column inside a block (layer block, then i, then k), so that a 3D cell's z
// neighbours are processed close to it while its r neighbours stay within a
// block; zb = 0 is the table order (layer by layer).  The lists of a zb are
// built on first use.
bool Engine::build_zblock(int zb) {
  if (tl_zb[zb][0]) return true;
  const int mask = (1 << zb) - 1;
  auto zone_key = [&](const int2 &t) {
    const AnnInfo &A = hann[t.x];
    return std::make_tuple(A.disk, A.k >> zb, A.i, A.k & mask);
  };
  auto face_key = [&](const int2 &t) {
    const AnnInfo &A = hann[hpairs[t.x].ann_in];
    return std::make_tuple(A.disk, A.k >> zb, A.i, A.k & mask);
  };
  std::vector<int2> zt = h_ztiles, ut = h_utiles, zf = h_zft;
  std::stable_sort(zt.begin(), zt.end(), [&](const int2 &a, const int2 &b) { return zone_key(a) < zone_key(b); });
  std::stable_sort(ut.begin(), ut.end(), [&](const int2 &a, const int2 &b) { return zone_key(a) < zone_key(b); });
  std::stable_sort(zf.begin(), zf.end(), [&](const int2 &a, const int2 &b) { return face_key(a) < face_key(b); });
  std::vector<int2> all(zt);
  all.insert(all.end(), ut.begin(), ut.end());
  all.insert(all.end(), zf.begin(), zf.end());
  if (!dalloc(&d_tiles_zb[zb], all.size(), "tiles") || !h2d(d_tiles_zb[zb], all.data(), all.size())) return false;
  tl_zb[zb][0] = d_tiles_zb[zb];
  tl_zb[zb][1] = d_tiles_zb[zb] + zt.size();
  tl_zb[zb][2] = d_tiles_zb[zb] + zt.size() + ut.size();
  return true;
}

void Engine::set_schedule(const Schedule &s0) {
  Schedule s = s0;
  if (!build_zblock(s.zb)) s.zb = 0;
  sched = s;
  plan.fuse = s.fuse;
  plan.ztiles = tl_zb[s.zb][0];
  plan.utiles = tl_zb[s.zb][1];
  plan.zftiles = tl_zb[s.zb][2];
  plan.stage_r = s.stage ? stage_r_max : 0;
  plan.stage_z = s.stage ? stage_z_max : 0;
  plan.face_threads = s.face_threads();
  if (gexec) {  // the next graph run captures the new launches
    hipGraphExecDestroy(gexec);
    hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }
  graph_cur = -1;
}

// Candidate k: Schedule::of(k) (xcd, largest_first, stage, fuse, zb, fw).
// Each runs on a scratch engine holding a copy of the batch's state (the
// batch itself does not move), graph-replayed for nsteps after two warm-up
// steps.  Schedules change only the order in which workgroups run, never the
// arithmetic, so every candidate gives the same bits.
int Engine::autotune(int nsteps, double *ms, int nms) {
  if (!ok) return -1;
  if (nsteps < 1) nsteps = 1;
  Engine t(ndisk, sims.data());
  if (!t.ok) return -1;
  for (int d = 0; d < ndisk; ++d) {
    std::vector<double> buf((size_t)NQ * nzones[d]);
    download(d, 0, buf.data());
    t.upload(d, 0, buf.data());
    for (int f : {5, 6, 7}) {
      download(d, f, buf.data());
      t.upload(d, f, buf.data());
    }
    t.put_scalars(d, get_scalars(d));
  }
  for (int k = 0; k < nms; ++k) ms[k] = -1.0;
  auto time_k = [&](int k) {
    t.set_schedule(Schedule::of(k));
    t.run(2, true);
    t.synchronize();
    t.run(nsteps, true);
    t.synchronize();
    const double v = t.last_ms / nsteps;
    if (k < nms) ms[k] = v;
    return v;
  };
  // the launch orders first (faces stored, table order), then the layer
  // blocks under the fastest order, then the fused face evaluations, then
  // the face builder's workgroup size
  double best = 1e300;
  int bestk = -1;
  auto consider = [&](int k) {
    if (!schedule_distinct(k)) return;  // the same launches as another candidate
    const double v = time_k(k);
    if (v < best) {
      best = v;
      bestk = k;
    }
  };
  for (int k = 0; k < 8; ++k) consider(k);
  if (bestk < 0) return -1;
  const int order = bestk;
  for (int zb = 1; zb < 8; ++zb) consider(order | zb << 5);
  const int blocked = bestk;
  for (int f = 1; f < 4; ++f) consider(blocked | f << 3);
  const int fused = bestk;
  for (int w = 1; w < 3; ++w) consider(fused | w << 8);
  set_schedule(Schedule::of(bestk));
  return bestk;
}

// false when candidate k launches exactly what a smaller candidate launches:
// LDS staging with nothing to stage, fused r faces with z faces present
bool Engine::schedule_distinct(int k) const {
  const Schedule s = Schedule::of(k);
  if (s.stage && !stage_r_max && !stage_z_max) return false;
  if ((s.fuse & 1) && n_zpairs) return false;
  if ((s.fuse & 2) && !n_zpairs) return false;
  if (s.zb && !n_zpairs) return false;  // 2D disks: every layer-block order is table order
  int nzmax = 1;
  for (const DiskParams &P : hdp) nzmax = P.NZ > nzmax ? P.NZ : nzmax;
  if (s.zb && (1 << (s.zb - 1)) >= nzmax) return false;  // blocks already span every layer at zb - 1
  if (s.fw == 3) return false;                            // (not a width)
  return true;
}

void Engine::synchronize() {
  hip_check(hipStreamSynchronize(st), "synchronize");
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, ev0, ev1) == hipSuccess) last_ms = ms;
}

int Engine::check_error() {
  int e = 0;
  hip_check(hipStreamSynchronize(st), "synchronize");
  hip_check(hipMemcpy(&e, d_err, sizeof(int), hipMemcpyDeviceToHost), "err");
  if (e == DERR_FACE_OVERFLOW) disco_set_error(30, "overlap faces exceed a pair's capacity (diverged mesh?)");
  if (e == DERR_VISC_VZ)
    disco_set_error(31, "viscosity needs v_z = 0 (the reference exits in riemann_visc_flux)");
  return e;
}

// The unit-level conserved state: allocated on the first unit-level call,
// and while not live it is what cell_calc_cons would leave (prim2cons of the
// primitives), derived on demand.
double *Engine::unit_cons() {
  if (!d_ucons) {
    if (!dalloc(&d_ucons, (size_t)NQ * Ntot, "unit cons") || !dalloc(&d_cool, (size_t)Ntot, "cool") ||
        !dalloc(&d_umdot, 2 * (size_t)(plan.n_ztiles ? plan.n_ztiles : 1), "unit sink"))
      return nullptr;
    hipMemsetAsync(d_cool, 0, sizeof(double) * Ntot, st);
    ucons_live = false;
  }
  if (!ucons_live) {
    launch_unit(mesh(), plan, st, NQ, UNIT_DERIVE, d_ucons, 0.0, 0.0, nullptr, nullptr, 0);
    ucons_live = true;
  }
  return d_ucons;
}

// the plugin's per-annulus single-cell rows of every disk, as fixed
// boundaries and damping would see them whatever the disk's boundary type
InitRow *Engine::unit_rows() {
  if (d_urows) return d_urows;
  std::vector<InitRow> all;
  for (int d = 0; d < ndisk; ++d) {
    std::vector<InitRow> rows;
    if (!disco_init_rows(sims[d], NQ, rows, nullptr, true)) return nullptr;
    all.insert(all.end(), rows.begin(), rows.end());
  }
  if (!dalloc(&d_urows, all.size(), "unit rows") || !h2d(d_urows, all.data(), all.size())) return nullptr;
  return d_urows;
}

int Engine::ledger_enable(bool on) {
  if (on && !d_led_torque) {
    if (!dalloc(&d_led_torque, (size_t)plan.n_utiles, "ledger")) return -1;
    hipMemsetAsync(d_led_torque, 0, sizeof(double) * plan.n_utiles, st);
  } else if (!on && d_led_torque) {
    hip_check(hipStreamSynchronize(st), "synchronize");
    hipFree(d_led_torque);
    d_led_torque = nullptr;
  }
  return 0;
}

// The last substep's boundary r-face increments and torque term of disk d:
// out[0, NQ + 3) what the faces between the inner ghost annulus and the first
// interior annulus added to their R cell (dt*dA*F[q], then dt*dA*Fvisc[SRR,
// LLL, TAU] when viscous), out[NQ + 3, 2 (NQ + 3)) the same for the faces
// between the last interior annulus and the outer ghosts, out[2 (NQ + 3)]
// the torque term sum dt dV rho F_phi r of the interior cells; sums in face
// and tile order
int Engine::ledger_last(int d, double *out) {
  const DiskParams &P = hdp[d];
  if (!d_led_torque || P.NglobalZ != 1 || P.NgMinR != 1) {
    disco_set_error(19, "ledger: enable it before the substep; 2D disks with one inner ghost annulus");
    return -1;
  }
  if (!hip_check(hipStreamSynchronize(st), "synchronize")) return -1;
  const int nv = NQ + 3;
  const int pairs[2] = {P.rpair_base, P.rpair_base + P.NR - 3};
  for (int side = 0; side < 2; ++side) {
    const PairInfo &Q = hpairs[pairs[side]];
    PairState ps;
    hip_check(hipMemcpy(&ps, d_pst + pairs[side], sizeof(ps), hipMemcpyDeviceToHost), "ledger D2H");
    std::vector<double> row(ps.count > 0 ? ps.count : 1);
    for (int v = 0; v < nv; ++v) {
      double s = 0.0;
      if (v < NQ || P.visc > 0.0) {
        hip_check(hipMemcpy(row.data(), d_Fface + (size_t)v * NFtot + Q.face_base, sizeof(double) * ps.count,
                            hipMemcpyDeviceToHost), "ledger D2H");
        for (int n = 0; n < ps.count; ++n) s += row[n];
      }
      out[side * nv + v] = s;
    }
  }
  std::vector<double> tq(plan.n_utiles);
  hip_check(hipMemcpy(tq.data(), d_led_torque, sizeof(double) * tq.size(), hipMemcpyDeviceToHost), "ledger D2H");
  double t = 0.0;
  for (int x = P.utile0; x < P.utile1; ++x) t += tq[x];
  out[2 * nv] = t;
  return check_error() ? -1 : 0;
}
