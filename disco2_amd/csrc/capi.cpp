// The reference's Cell / Face / TimeStep class API on top of the Engine,
// the batch API and error reporting.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/disco.h"
#include "engine.h"
#include "initial_data.h"

bool diag_h5_write(const char *fname, const double *data, size_t n0, size_t n1);  // diagnostics.cpp
#include "h5.h"

// ------------------------------------------------------------------ errors
static int g_err = 0;
static std::string g_msg;
static std::mutex g_mu;
void disco_set_error(int code, const std::string &msg) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_err == 0) {
    g_err = code;
    g_msg = msg;
  }
}
extern "C" {
int disco_last_error(void) { return g_err; }
const char *disco_last_error_msg(void) { return g_msg.c_str(); }
void disco_clear_error(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_err = 0;
  g_msg.clear();
}
const char *disco_version(void) { return "disco2_amd 0.1 (gfx950, fp64)"; }
}

// ------------------------------------------------------------------ helpers
static CellsHandle *H(Cell ***c) { return reinterpret_cast<CellsHandle *>(c); }

static void mirror_pull(CellsHandle *h) {
  if (!h->stale) return;
  Engine *e = h->eng;
  e->download(0, 0, h->prim.data());
  e->download(0, 3, h->grad.data());
  e->download(0, 4, h->gradp.data());
  e->download(0, 5, h->tiph.data());
  e->download(0, 7, h->dphi.data());
  e->download(0, 8, h->wiph.data());
  h->stale = false;
}
static void mirror_push(CellsHandle *h) {
  if (h->dirty_prim) h->eng->upload(0, 0, h->prim.data());
  if (h->dirty_tiph) h->eng->upload(0, 5, h->tiph.data());
  h->dirty_prim = h->dirty_tiph = false;
}
static long long zone_of(CellsHandle *h, int i, int j, int k) {
  const AnnInfo &A = h->eng->hann[k * h->eng->hdp[0].NR + i];
  return A.zbase + j;
}
static void scalars_from_host(Engine *e, TimeStep *ts, GravMass *gm) {
  DiskScalars s = e->get_scalars(0);
  if (ts) {
    s.t = ts->t;
    s.dt = ts->dt;
    s.RK = ts->RK;
  }
  if (gm)
    for (int p = 0; p < 2; ++p) {
      s.gm[p].r = gm[p].r; s.gm[p].phi = gm[p].phi; s.gm[p].M = gm[p].M; s.gm[p].omega = gm[p].omega;
      s.gm[p].RK_r = gm[p].RK_r; s.gm[p].RK_phi = gm[p].RK_phi; s.gm[p].RK_M = gm[p].RK_M;
      s.gm[p].RK_omega = gm[p].RK_omega; s.gm[p].Mdot = gm[p].Mdot; s.gm[p].Macc = gm[p].Macc;
    }
  e->put_scalars(0, s);
  launch_scalar(e->mesh(), e->plan, e->st, SC_MASS);
}
static void scalars_to_host(Engine *e, TimeStep *ts, GravMass *gm) {
  DiskScalars s = e->get_scalars(0);
  if (ts) {
    ts->t = s.t;
    ts->dt = s.dt;
    ts->RK = s.RK;
  }
  if (gm)
    for (int p = 0; p < 2; ++p) {
      gm[p].r = s.gm[p].r; gm[p].phi = s.gm[p].phi; gm[p].M = s.gm[p].M; gm[p].omega = s.gm[p].omega;
      gm[p].RK_r = s.gm[p].RK_r; gm[p].RK_phi = s.gm[p].RK_phi; gm[p].RK_M = s.gm[p].RK_M;
      gm[p].RK_omega = s.gm[p].RK_omega; gm[p].Mdot = s.gm[p].Mdot; gm[p].Macc = s.gm[p].Macc;
    }
}

extern "C" {
// ------------------------------------------------------------------ Cell
// cell_create, src/Cell/cell_create_destroy.c:23-95: zero primitives, one
// random phi offset per annulus (srand(666) sequence), dphi = 2 pi / N_p;
// the edges are also wrapped and z-synced as src/main.c:80-85 does next.
Cell ***cell_create(Sim *s, MPIsetup *) {
  if (!s->allocated) {
    disco_set_error(10, "cell_create before sim_alloc_arr");
    return nullptr;
  }
  Engine *e = new Engine(1, &s);
  if (!e->ok) {
    delete e;
    return nullptr;
  }
  CellsHandle *h = new CellsHandle();
  h->eng = e;
  h->sim = s;
  h->nz = e->nzones[0];
  h->NQ = e->NQ;
  h->prim.assign((size_t)h->NQ * h->nz, 0.0);
  h->grad.assign((size_t)h->NQ * h->nz, 0.0);
  h->gradp.assign((size_t)h->NQ * h->nz, 0.0);
  h->tiph.assign(h->nz, 0.0);
  h->dphi.assign(h->nz, 0.0);
  h->wiph.assign(h->nz, 0.0);
  h->stale = true;
  return reinterpret_cast<Cell ***>(h);
}

void cell_destroy(Cell ***c, Sim *) {
  CellsHandle *h = H(c);
  if (!h) return;
  delete h->eng;
  delete[] h->refs;
  delete h;
}

long long cell_nzones(Cell ***c) { return H(c)->nz; }

Cell *cell_single(Cell ***c, int i, int j, int k) {
  CellsHandle *h = H(c);
  if (!h->refs) {
    h->refs = new CellRef[h->nz];
    for (long long z = 0; z < h->nz; ++z) h->refs[z] = CellRef{h, z};
  }
  return reinterpret_cast<Cell *>(&h->refs[zone_of(h, i, j, k)]);
}
long long cell_zone_index(Cell *c) { return reinterpret_cast<CellRef *>(c)->zone; }

static CellRef *R(Cell *c) {
  CellRef *r = reinterpret_cast<CellRef *>(c);
  if (!r->own) mirror_pull(r->h);
  return r;
}
// a cell of its own keeps NQ primitives, then tiph, dphi, wiph; it has no
// slopes (the reference's cell_single_create allocates none)
double cell_prim(Cell *c, int q) {
  CellRef *r = R(c);
  return r->own ? r->own[q] : r->h->prim[q * r->h->nz + r->zone];
}
double cell_grad(Cell *c, int q) { CellRef *r = R(c); return r->own ? 0.0 : r->h->grad[q * r->h->nz + r->zone]; }
double cell_gradp(Cell *c, int q) { CellRef *r = R(c); return r->own ? 0.0 : r->h->gradp[q * r->h->nz + r->zone]; }
double cell_tiph(Cell *c) { CellRef *r = R(c); return r->own ? r->own[r->NQ] : r->h->tiph[r->zone]; }
double cell_dphi(Cell *c) { CellRef *r = R(c); return r->own ? r->own[r->NQ + 1] : r->h->dphi[r->zone]; }
double cell_wiph(Cell *c) { CellRef *r = R(c); return r->own ? r->own[r->NQ + 2] : r->h->wiph[r->zone]; }

// ------------------------------------------------------------------ initial data
// The plugins of src/Cell/initial_data/*.c run on the host over the cell
// mirror (initial_data.cpp); the device copy follows before the next launch.
static void cell_init_type(Cell ***c, Sim *s, GravMass *g, int type) {
  CellsHandle *h = H(c);
  mirror_pull(h);
  if (disco_cell_init(s, type, g, h->tiph.data(), h->dphi.data(), h->prim.data(), h->nz, h->NQ))
    h->dirty_prim = true;
}
static void cell_single_init_type(Cell *cell, Sim *s, int type, int i) {
  CellRef *r = R(cell);
  if (r->own) {  // a cell of its own (cell_single_create)
    double o[DISCO_MAXQ];
    for (int q = 0; q < r->NQ; ++q) o[q] = r->own[q];
    if (disco_single_init(s, type, i, r->own[r->NQ], r->own[r->NQ + 1], o, r->NQ) <= 0) return;
    for (int q = 0; q < r->NQ; ++q) r->own[q] = o[q];
    return;
  }
  CellsHandle *h = r->h;
  double o[DISCO_MAXQ];
  for (int q = 0; q < h->NQ; ++q) o[q] = h->prim[q * h->nz + r->zone];
  if (disco_single_init(s, type, i, h->tiph[r->zone], h->dphi[r->zone], o, h->NQ) <= 0) return;
  for (int q = 0; q < h->NQ; ++q) h->prim[q * h->nz + r->zone] = o[q];
  h->dirty_prim = true;
}
#define DISCO_PLUGIN(name, type)                                                                        \
  void cell_init_##name(Cell ***c, Sim *s, GravMass *g, MPIsetup *) { cell_init_type(c, s, g, type); } \
  void cell_single_init_##name(Cell *c, Sim *s, GravMass *, int i, int, int) { cell_single_init_type(c, s, type, i); }
DISCO_PLUGIN(shear, ID_SHEAR)
DISCO_PLUGIN(vortex, ID_VORTEX)
DISCO_PLUGIN(SStest, ID_SSTEST)
DISCO_PLUGIN(torus, ID_TORUS)
DISCO_PLUGIN(milos_macfadyen, ID_MILOS_MACFADYEN)
DISCO_PLUGIN(rad_dom, ID_RAD_DOM)
DISCO_PLUGIN(middle, ID_MIDDLE)
#undef DISCO_PLUGIN

// src/Cell/cell_init_function_ptr.c:6-46; NULL (error set) where the
// reference prints an error and exits
void (*cell_init_ptr(Sim *s))(Cell ***, Sim *, GravMass *, MPIsetup *) {
  switch (s->InitialDataType) {
    case ID_SHEAR: return &cell_init_shear;
    case ID_VORTEX: return &cell_init_vortex;
    case ID_TORUS: return &cell_init_torus;
    case ID_MILOS_MACFADYEN: return &cell_init_milos_macfadyen;
    case ID_RAD_DOM: return &cell_init_rad_dom;
    case ID_MIDDLE: return &cell_init_middle;
    case ID_SSTEST: return &cell_init_SStest;
  }
  disco_set_error(8, "not a valid global initial data routine choice");
  return nullptr;
}
void (*cell_single_init_ptr(Sim *s))(Cell *, Sim *, GravMass *, int, int, int) {
  switch (s->InitialDataType) {
    case ID_SHEAR: return &cell_single_init_shear;
    case ID_VORTEX: return &cell_single_init_vortex;
    case ID_TORUS: return &cell_single_init_torus;
    case ID_MILOS_MACFADYEN: return &cell_single_init_milos_macfadyen;
    case ID_RAD_DOM: return &cell_single_init_rad_dom;
    case ID_MIDDLE: return &cell_single_init_middle;
    case ID_SSTEST: return &cell_single_init_SStest;
  }
  disco_set_error(8, "not a valid single point initial data routine choice");
  return nullptr;
}

void cell_set_prim(Cell ***c, int i, int j, int k, int q, double v) {  // src/Cell/cell_modify_data.c:13-15
  CellsHandle *h = H(c);
  mirror_pull(h);
  h->prim[q * h->nz + zone_of(h, i, j, k)] = v;
  h->dirty_prim = true;
}
void cell_set_tiph(Cell ***c, int i, int j, int k, double v) {  // src/Cell/cell_modify_data.c:16-18
  CellsHandle *h = H(c);
  mirror_pull(h);
  h->tiph[zone_of(h, i, j, k)] = v;
  h->dirty_tiph = true;
}

void cell_upload(Cell ***c, int field, const double *src) {
  CellsHandle *h = H(c);
  mirror_push(h);
  h->eng->upload(0, field, src);
  h->stale = true;
}
void cell_download(Cell ***c, int field, double *dst) {
  CellsHandle *h = H(c);
  mirror_push(h);
  h->eng->download(0, field, dst);
}

// src/Cell/cell_misc.c:12-26
void cell_clean_pi(Cell ***c, Sim *) {
  CellsHandle *h = H(c);
  mirror_pull(h);
  for (auto &phi : h->tiph) {
    while (phi > 2.0 * M_PI) phi -= 2.0 * M_PI;
    while (phi < 0.0) phi += 2.0 * M_PI;
  }
  h->dirty_tiph = true;
  mirror_push(h);
}
void cell_syncproc_r(Cell ***, Sim *, MPIsetup *) {}  // one rank: no r neighbours
// src/Cell/cell_sync.c:156-205 on one rank: periodic z copy of prim, RKcons, tiph
void cell_syncproc_z(Cell ***c, Sim *s, MPIsetup *) {
  CellsHandle *h = H(c);
  if (s->N_global[1] <= 1) return;
  mirror_pull(h);
  mirror_push(h);
  Engine *e = h->eng;
  const DiskParams &P = e->hdp[0];
  long long nz = h->nz, NL = P.NL;
  std::vector<double> rk((size_t)h->NQ * nz);
  e->download(0, 2, rk.data());
  int NZ = P.NZ, kNm2g = NZ - 2 * P.NgMaxZ, kNmg = NZ - P.NgMaxZ, kg = P.NgMinZ;
  auto cp = [&](int kd, int ks) {
    for (long long x = 0; x < NL; ++x) {
      for (int q = 0; q < h->NQ; ++q) {
        h->prim[q * nz + kd * NL + x] = h->prim[q * nz + ks * NL + x];
        rk[q * nz + kd * NL + x] = rk[q * nz + ks * NL + x];
      }
      h->tiph[kd * NL + x] = h->tiph[ks * NL + x];
    }
  };
  for (int d = 0; d < P.NgMaxZ; ++d) cp(kNmg + d, kg + d);
  for (int d = 0; d < P.NgMinZ; ++d) cp(d, kNm2g + d);
  e->upload(0, 0, h->prim.data());
  e->upload(0, 5, h->tiph.data());
  e->upload(0, 2, rk.data());
}
// src/Cell/cell_conversion.c:35-55: the conserved state is prim2cons of the
// primitives again (derived on the device when a unit-level call needs it)
void cell_calc_cons(Cell ***c, Sim *) {
  mirror_push(H(c));
  H(c)->eng->ucons_live = false;
}

// src/Cell/cell_conversion.c:11-33
void cell_prim2cons(double *prim, double *cons, double r, double dV, Sim *s) {
  double G = s->GAMMALAW;
  double rho = prim[RHO], Pp = prim[PPP], vr = prim[URR], vp = prim[UPP] * r, vz = prim[UZZ];
  double v2 = vr * vr + vp * vp + vz * vz;
  double rhoe = Pp / (G - 1.);
  cons[DDD] = rho * dV;
  cons[SRR] = rho * vr * dV;
  cons[LLL] = r * rho * vp * dV;
  cons[SZZ] = rho * vz * dV;
  cons[TAU] = (.5 * rho * v2 + rhoe) * dV;
  for (int q = s->NUM_C; q < s->NUM_Q; ++q) cons[q] = prim[q] * cons[DDD];
}
// src/Cell/cell_conversion.c:57-103
void cell_cons2prim(double *cons, double *prim, double r, double dV, Sim *s) {
  double G = s->GAMMALAW;
  double rho = cons[DDD] / dV;
  if (rho < s->RHO_FLOOR) rho = s->RHO_FLOOR;
  double Sr = cons[SRR] / dV, Sp = cons[LLL] / dV / r, Sz = cons[SZZ] / dV, E = cons[TAU] / dV;
  double vr = Sr / rho, vp = Sp / rho, vz = Sz / rho;
  double KE = .5 * (Sr * vr + Sp * vp + Sz * vz);
  double v_magnitude = sqrt(2. * KE / rho);
  double rhoe = E - KE;
  double Pp = (G - 1.) * rhoe;
  if (Pp < s->CS_FLOOR * s->CS_FLOOR * rho / G) Pp = s->CS_FLOOR * s->CS_FLOOR * rho / G;
  if (Pp > s->CS_CAP * s->CS_CAP * rho / G) Pp = s->CS_CAP * s->CS_CAP * rho / G;
  if (v_magnitude > s->VEL_CAP) {
    vr = vr * s->VEL_CAP / v_magnitude;
    vp = vp * s->VEL_CAP / v_magnitude;
    vz = vz * s->VEL_CAP / v_magnitude;
  }
  prim[RHO] = rho;
  prim[PPP] = Pp;
  prim[URR] = vr;
  prim[UPP] = vp / r;
  prim[UZZ] = vz;
  for (int q = s->NUM_C; q < s->NUM_Q; ++q) prim[q] = cons[q] / cons[DDD];
}
// cell_calc_prim, src/Cell/cell_conversion.c:105-122, from the unit-level
// conserved state (prim2cons of the primitives unless a unit-level call
// changed it)
void cell_calc_prim(Cell ***c, Sim *) {
  CellsHandle *h = H(c);
  mirror_push(h);
  Engine *e = h->eng;
  double *uc = e->unit_cons();
  if (!uc) return;
  launch_unit(e->mesh(), e->plan, e->st, e->NQ, UNIT_CALC_PRIM, uc, 0.0, 0.0, nullptr, nullptr, 0);
  e->check_error();
  h->stale = true;
}

// src/Cell/cell_misc.c:134-143.  The next step recomputes wiph (the engine's
// cached mesh velocity is dropped), as the reference's timestep_rk2 always
// runs cell_set_w after this.
void cell_clear_w(Cell ***c, Sim *) {
  CellsHandle *h = H(c);
  mirror_push(h);
  hipMemsetAsync(h->eng->d_wiph, 0, sizeof(double) * h->eng->nann, h->eng->st);
  h->eng->w_ready = false;
  h->stale = true;
}
void cell_set_w(Cell ***c, Sim *) {
  CellsHandle *h = H(c);
  mirror_push(h);
  launch_set_w(h->eng->mesh(), h->eng->plan, h->eng->st);
  h->stale = true;
}
double cell_mindt(Cell ***c, Sim *, GravMass *gm) {  // src/Cell/cell_mindt.c:32-100
  CellsHandle *h = H(c);
  mirror_push(h);
  Engine *e = h->eng;
  scalars_from_host(e, nullptr, gm);
  launch_scalar(e->mesh(), e->plan, e->st, SC_RESET_DT | SC_MASS);
  launch_mindt(e->mesh(), e->plan, e->st, e->NQ);
  launch_scalar(e->mesh(), e->plan, e->st, SC_DT_FOLD);
  DiskScalars s = e->get_scalars(0);
  double dt;
  memcpy(&dt, &s.dtmin_bits, 8);
  return dt;
}
void cell_plm_p(Cell ***c, Sim *) {
  CellsHandle *h = H(c);
  mirror_push(h);
  launch_plm(h->eng->mesh(), h->eng->plan, h->eng->st, h->eng->NQ, -1);
  h->stale = true;
}
void cell_plm_rz(Cell ***c, Sim *, Face *, TimeStep *, MPIsetup *, int direction) {
  CellsHandle *h = H(c);
  mirror_push(h);
  launch_plm(h->eng->mesh(), h->eng->plan, h->eng->st, h->eng->NQ, direction);
  h->stale = true;
}

// ------------------------------------------------------------------ Face
// face_create, src/Face/face_create_destroy.c:130-193, run on the device;
// the host copy lists the faces in the reference's order (i outer, k inner
// for r; k outer, i inner for z) and fills the TimeStep offsets nri / nzk.
Face *face_create(Cell ***c, Sim *, TimeStep *ts, int direction) {
  CellsHandle *h = H(c);
  mirror_push(h);
  Engine *e = h->eng;
  launch_faces(e->mesh(), e->plan, e->st, e->NQ, direction);
  if (e->check_error()) return nullptr;
  const DiskParams &P = e->hdp[0];
  if (direction == 1 && P.NglobalZ <= 1) return nullptr;
  std::vector<PairState> pst(e->hpairs.size());
  hipMemcpy(pst.data(), e->d_pst, pst.size() * sizeof(PairState), hipMemcpyDeviceToHost);
  std::vector<int> fL(e->NFtot), fR(e->NFtot);
  std::vector<double> fd(e->NFtot), fc(e->NFtot);
  hipMemcpy(fL.data(), e->d_fL, fL.size() * sizeof(int), hipMemcpyDeviceToHost);
  hipMemcpy(fR.data(), e->d_fR, fR.size() * sizeof(int), hipMemcpyDeviceToHost);
  hipMemcpy(fd.data(), e->d_fdphi, fd.size() * sizeof(double), hipMemcpyDeviceToHost);
  hipMemcpy(fc.data(), e->d_fcm, fc.size() * sizeof(double), hipMemcpyDeviceToHost);
  Face *F = new Face();
  F->h = h;
  F->dir = direction;
  int NR = P.NR, NZ = P.NZ;
  auto append = [&](int p) {
    const PairInfo &Q = e->hpairs[p];
    for (int n = 0; n < pst[p].count; ++n) {
      long long g = Q.face_base + n;
      F->L.push_back(fL[g]);
      F->R.push_back(fR[g]);
      F->r.push_back(Q.r);
      F->dphi.push_back(fd[g]);
      F->dA.push_back(Q.r * fd[g] * Q.deltaPerp);
      F->deltaL.push_back(Q.deltaL);
      F->deltaR.push_back(Q.deltaR);
      F->cm.push_back(fc[g]);
    }
  };
  if (direction == 0) {
    if (ts) ts->nri.assign(NR, 0);
    for (int i = 0; i < NR - 1; ++i) {
      if (ts) ts->nri[i] = (int)F->L.size();
      for (int k = 0; k < NZ; ++k) append(P.rpair_base + k * (NR - 1) + i);
    }
    if (ts) ts->nri[NR - 1] = (int)F->L.size();
  } else {
    if (ts) ts->nzk.assign(NZ, 0);
    for (int k = 0; k < NZ - 1; ++k) {
      if (ts) ts->nzk[k] = (int)F->L.size();
      for (int i = 0; i < NR; ++i) append(P.zpair_base + k * NR + i);
    }
    if (ts) ts->nzk[NZ - 1] = (int)F->L.size();
  }
  return F;
}
void face_destroy(Face *f) { delete f; }
int face_count(Face *f) { return f ? (int)f->L.size() : 0; }
static Cell *ref_of(CellsHandle *h, long long z) {
  if (!h->refs) {
    h->refs = new CellRef[h->nz];
    for (long long x = 0; x < h->nz; ++x) h->refs[x] = CellRef{h, x};
  }
  return reinterpret_cast<Cell *>(&h->refs[z]);
}
Cell *face_L_pointer(Face *f, int n) { return ref_of(f->h, f->L[n]); }
Cell *face_R_pointer(Face *f, int n) { return ref_of(f->h, f->R[n]); }
double face_deltaL(Face *f, int n) { return f->deltaL[n]; }
double face_deltaR(Face *f, int n) { return f->deltaR[n]; }
double face_cm(Face *f, int n) { return f->cm[n]; }
double face_dA(Face *f, int n) { return f->dA[n]; }
double face_r(Face *f, int n) { return f->r[n]; }

// ------------------------------------------------------------------ TimeStep
TimeStep *timestep_create(Sim *s) {  // src/TimeStep/timestep_create_destroy.c:7-13
  TimeStep *t = new TimeStep();
  t->nri.assign(sim_N(s, 0), 0);
  t->nzk.assign(sim_N(s, 1), 0);
  return t;
}
void timestep_destroy(TimeStep *t) { delete t; }
double timestep_get_t(TimeStep *t) { return t->t; }
double timestep_dt(TimeStep *t) { return t->dt; }
void timestep_set_t(TimeStep *t, double v) { t->t = v; }
void timestep_update_t(TimeStep *t) { t->t += t->dt; }
void timestep_set_RK(TimeStep *t, double RK) { t->RK = RK; }
void timestep_set_nri(TimeStep *t, int i, int n) { t->nri[i] = n; }
void timestep_set_nzk(TimeStep *t, int k, int n) { t->nzk[k] = n; }
int timestep_n(TimeStep *t, int index, int direction) {  // src/TimeStep/timestep_access.c:30-39
  if (direction == 0) return t->nri[index];
  if (direction == 1) return t->nzk[index];
  disco_set_error(4, "timestep_n: bad direction");
  return 0;
}
// src/TimeStep/timestep_set.c:16-21
void timestep_set_dt(TimeStep *ts, Cell ***c, Sim *s, GravMass *gm) {
  ts->dt = cell_mindt(c, s, gm);
  if (ts->t + ts->dt > s->T_MAX) ts->dt = s->T_MAX - ts->t;
}
// One RK2 step on the device; t, dt and both masses come back to the host
// objects at return, as the reference's loop reads t after every step.
void timestep_rk2(TimeStep *ts, Sim *, Cell ***c, GravMass *gm, MPIsetup *) {
  CellsHandle *h = H(c);
  mirror_push(h);
  Engine *e = h->eng;
  scalars_from_host(e, ts, gm);
  e->enqueue_rk2();
  e->ucons_live = false;  // cell_calc_cons ends every substep
  scalars_to_host(e, ts, gm);
  e->check_error();
  h->stale = true;
}
void timestep_forward_euler(TimeStep *ts, Sim *, Cell ***c, GravMass *gm, MPIsetup *) {
  CellsHandle *h = H(c);
  mirror_push(h);
  Engine *e = h->eng;
  scalars_from_host(e, ts, gm);
  e->enqueue_euler();
  e->ucons_live = false;  // cell_calc_cons ends every substep
  scalars_to_host(e, ts, gm);
  e->check_error();
  h->stale = true;
}
// One substep with the RK snapshot already on the device (the reference
// takes it with cell_copy before the first substep).
void timestep_substep(TimeStep *ts, Cell ***c, Sim *, GravMass *gm, MPIsetup *, double fac) {
  CellsHandle *h = H(c);
  mirror_push(h);
  Engine *e = h->eng;
  scalars_from_host(e, ts, gm);
  e->enqueue_substep(fac, 0, 1, SC_SUB_BEGIN, SC_AFTER_UPDATE);
  e->ucons_live = false;  // cell_calc_cons ends every substep
  scalars_to_host(e, ts, gm);
  e->check_error();
  h->stale = true;
}

// ------------------------------------------------------------------ batch
struct DiscoBatch {
  Engine *e;
};
DiscoBatch *disco_batch_create(int n, Sim **sims) {
  for (int d = 0; d < n; ++d)
    if (!sims[d]->allocated) {
      disco_set_error(10, "batch disk without sim_alloc_arr");
      return nullptr;
    }
  Engine *e = new Engine(n, sims);
  if (!e->ok) {
    delete e;
    return nullptr;
  }
  return new DiscoBatch{e};
}
void disco_batch_destroy(DiscoBatch *b) {
  if (!b) return;
  delete b->e;
  delete b;
}
long long disco_batch_nzones(DiscoBatch *b, int d) { return b->e->nzones[d]; }
void disco_batch_upload(DiscoBatch *b, int d, int field, const double *src) { b->e->upload(d, field, src); }
void disco_batch_download(DiscoBatch *b, int d, int field, double *dst) { b->e->download(d, field, dst); }
void disco_batch_init(DiscoBatch *b) { b->e->init_state(); }
void disco_batch_run(DiscoBatch *b, int nsteps, int graph) {
  b->e->run(nsteps, graph != 0);
}
void disco_batch_synchronize(DiscoBatch *b) {
  b->e->synchronize();
  b->e->check_error();
}
double disco_batch_last_ms(DiscoBatch *b) { return b->e->last_ms; }
int disco_batch_torque_loss(DiscoBatch *b, int nsteps, double *loss) { return b->e->torque_loss(nsteps, loss); }
int disco_batch_torque_loss_noalloc(DiscoBatch *b, int nsteps, double *loss) {
  return b->e->torque_loss_noalloc(nsteps, loss);
}
int disco_batch_adjoint(DiscoBatch *b, int nsteps, int every, double *loss, double *g_alpha, double *g_q,
                        double *g_prim0) {
  return b->e->adjoint(nsteps, every, loss, g_alpha, g_q, g_prim0);
}

// Eager steps with every kernel launch bracketed by HIP events on the
// batch's stream; per kernel name: total ms and launch count.
int disco_batch_profile(DiscoBatch *b, int nsteps, char *names, double *ms, int *counts, int maxk) {
  Engine *e = b->e;
  int cap = nsteps * 64 + 16;
  std::vector<KMark> marks(cap);
  for (auto &m : marks) {
    hipEventCreate(&m.a);
    hipEventCreate(&m.b);
  }
  KProf prof{marks.data(), 0, cap};
  e->synchronize();
  g_kprof = &prof;
  for (int s = 0; s < nsteps; ++s) e->enqueue_rk2();
  g_kprof = nullptr;
  hip_check(hipStreamSynchronize(e->st), "profile");
  int nk = 0;
  std::vector<std::string> keys;
  for (int i = 0; i < prof.n; ++i) {
    float t = 0.f;
    hipEventElapsedTime(&t, marks[i].a, marks[i].b);
    std::string nm(marks[i].name);
    int k = (int)(std::find(keys.begin(), keys.end(), nm) - keys.begin());
    if (k == (int)keys.size()) {
      if (nk >= maxk) continue;
      keys.push_back(nm);
      ms[nk] = 0.0;
      counts[nk] = 0;
      strncpy(names + 32 * nk, nm.c_str(), 31);
      names[32 * nk + 31] = 0;
      ++nk;
    }
    ms[k] += t;
    counts[k] += 1;
  }
  for (auto &m : marks) {
    hipEventDestroy(m.a);
    hipEventDestroy(m.b);
  }
  e->check_error();
  return nk;
}

// A plain device copy of n doubles (known HBM traffic: 8n read + 8n write),
// used to calibrate hardware byte counters; returns the copy time in ms.
double disco_calibrate_copy(long long n) {
  double *a = nullptr, *b = nullptr;
  if (!hip_check(hipMalloc((void **)&a, n * sizeof(double)), "calib") ||
      !hip_check(hipMalloc((void **)&b, n * sizeof(double)), "calib"))
    return -1.0;
  hipMemset(a, 0, n * sizeof(double));
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  launch_copy(b, a, n, 0);
  hipEventRecord(e0, 0);
  launch_copy(b, a, n, 0);
  hipEventRecord(e1, 0);
  hipEventSynchronize(e1);
  float t = 0.f;
  hipEventElapsedTime(&t, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(a);
  hipFree(b);
  return t;
}
void disco_batch_scalars(DiscoBatch *b, int d, double *o) {
  DiskScalars s = b->e->get_scalars(d);
  o[0] = s.t;
  o[1] = s.dt;
  o[2] = s.RK;
  for (int p = 0; p < 2; ++p) {
    double *g = o + 3 + 10 * p;
    g[0] = s.gm[p].r; g[1] = s.gm[p].phi; g[2] = s.gm[p].M; g[3] = s.gm[p].omega;
    g[4] = s.gm[p].RK_r; g[5] = s.gm[p].RK_phi; g[6] = s.gm[p].RK_M; g[7] = s.gm[p].RK_omega;
    g[8] = s.gm[p].Mdot; g[9] = s.gm[p].Macc;
  }
}

// ------------------------------------------------------------------ diagnostics
// src/Diagnostics, one rank: the measurement on the device
// (Engine::diag_measure); schedule, running sums and files as
// diagnostics_create_destroy.c:9-118, diagnostics_set.c:13-303 and
// diagnostics_print.c:11-92.
}  // extern "C"
struct Diagnostics {
  Engine *eng = nullptr;  // of the last measurement
  int nrg = 0, neq = 0;
  std::vector<double> vec, scal, equat;
  double toutprev = 0, toutprev_dump = 0;
  double tdiag_measure = 0, dtdiag_measure = 0, tdiag_dump = 0, dtdiag_dump = 0;
  double rz[4] = {0, 0, 0, 0};
};
extern "C" {

Diagnostics *diagnostics_create(Sim *sim, TimeStep *ts, MPIsetup *) {
  Diagnostics *d = new Diagnostics;
  const int imin = sim_Nghost_min(sim, 0), imax = sim_N(sim, 0) - sim_Nghost_max(sim, 0);
  const int kmin = sim_Nghost_min(sim, 1), kmax = sim_N(sim, 1) - sim_Nghost_max(sim, 1);
  for (int k = kmin; k < kmax; ++k) {
    const double zp = sim_FacePos(sim, k, 1), zm = sim_FacePos(sim, k - 1, 1), z = 0.5 * (zm + zp);
    for (int i = imin; i < imax; ++i)
      for (int j = 0; j < sim_N_p(sim, i); ++j)
        if ((fabs(zp) < 0.0000001) || (fabs(z) < 0.0000001)) ++d->neq;
  }
  d->nrg = sim_N_global(sim, 0);
  d->vec.assign((size_t)d->nrg * 17, 0.0);
  d->scal.assign(16, 0.0);
  d->equat.assign((size_t)d->neq * 18, 0.0);
  const double t = timestep_get_t(ts);
  d->toutprev = d->toutprev_dump = t;
  d->dtdiag_measure = sim_get_T_MAX(sim) / sim_NUM_DIAG_MEASURE(sim);
  d->tdiag_measure = d->dtdiag_measure;
  d->dtdiag_dump = sim_get_T_MAX(sim) / sim_NUM_DIAG_DUMP(sim);
  d->tdiag_dump = d->dtdiag_dump;
  while (d->tdiag_measure < t) d->tdiag_measure += d->dtdiag_measure;
  while (d->tdiag_dump < t) d->tdiag_dump += d->dtdiag_dump;
  d->rz[0] = sim_MIN(sim, 0);
  d->rz[1] = sim_MAX(sim, 0);
  d->rz[2] = sim_MIN(sim, 1);
  d->rz[3] = sim_MAX(sim, 1);
  return d;
}

void diagnostics_destroy(Diagnostics *d, Sim *) { delete d; }

double diagnostics_tdiag_measure(Diagnostics *d) { return d->tdiag_measure; }
double diagnostics_tdiag_dump(Diagnostics *d) { return d->tdiag_dump; }

void diagnostics_set(Diagnostics *d, Cell ***c, Sim *, TimeStep *ts, MPIsetup *, GravMass *gm) {
  const double t = timestep_get_t(ts);
  if (!(t > d->tdiag_measure)) return;
  CellsHandle *h = H(c);
  mirror_push(h);
  Engine *e = h->eng;
  scalars_from_host(e, ts, gm);
  std::vector<double> v(d->vec.size()), s(16), md(12);
  if (e->diag_measure(0, t - d->toutprev, d->rz, v.data(), s.data(), nullptr, md.data()) != 0) return;
  d->eng = e;  // the equatorial cells stay on the device until a dump
  for (int p = 0; p < 2; ++p) {  // the masses' accretion as the host GravMass holds it
    md[4 + p] = gm[p].Mdot;
    md[6 + p] = gm[p].Macc;
  }
  if (FILE *f = fopen("DiagMdot.dat", "a")) {
    fprintf(f, "%e", t);
    for (double x : md) fprintf(f, " %e", x);
    fprintf(f, "\n");
    fclose(f);
  }
  for (size_t x = 0; x < v.size(); ++x) d->vec[x] += v[x];
  for (int n = 0; n < 16; ++n) d->scal[n] += s[n];
  d->toutprev = t;
  d->tdiag_measure += d->dtdiag_measure;
}

void diagnostics_print(Diagnostics *d, TimeStep *ts, Sim *, MPIsetup *) {
  const double t = timestep_get_t(ts);
  if (!(t > d->tdiag_dump)) return;
  const double dt_dump = t - d->toutprev_dump;
  char fe[256], fv[256];
  snprintf(fe, sizeof fe, "DiagEquat_%010.2f.h5", t);
  snprintf(fv, sizeof fv, "DiagVector_%010.2f.dat", t);
  FILE *vf = fopen(fv, "w"), *sf = fopen("DiagScalar.dat", "a");
  if (vf) {
    for (int i = 0; i < d->nrg; ++i) {
      for (int n = 0; n < 17; ++n) fprintf(vf, "%e ", d->vec[(size_t)i * 17 + n] / dt_dump);
      fprintf(vf, "\n");
    }
    fclose(vf);
  }
  if (sf) {
    fprintf(sf, "%e ", t);
    for (int n = 0; n < 16; ++n) fprintf(sf, "%e ", d->scal[n] / dt_dump);
    fprintf(sf, "\n");
    fclose(sf);
  }
  if (d->eng && d->eng->diag_equat(d->equat.data()) != 0) return;
  diag_h5_write(fe, d->equat.data(), (size_t)d->neq, 18);
  d->toutprev_dump = t;
  std::fill(d->vec.begin(), d->vec.end(), 0.0);
  std::fill(d->scal.begin(), d->scal.end(), 0.0);
  d->tdiag_dump += d->dtdiag_dump;
}

int disco_batch_diag_sizes(DiscoBatch *b, int d, int *nrg, int *neq) {
  if (d < 0 || d >= b->e->ndisk) {
    disco_set_error(14, "diagnostics: no such disk");
    return -1;
  }
  return b->e->diag_sizes(d, nrg, neq);
}
int disco_batch_diag_measure(DiscoBatch *b, int d, double dtout, const double *rz, double *vec, double *scal,
                             double *equat, double *mdot) {
  return b->e->diag_measure(d, dtout, rz, vec, scal, equat, mdot);
}
int disco_batch_diag_equat(DiscoBatch *b, double *equat) { return b->e->diag_equat(equat); }
void disco_batch_set_scalars(DiscoBatch *b, int d, const double *in) {
  Engine *e = b->e;
  DiskScalars s = e->get_scalars(d);
  s.t = in[0];
  s.dt = in[1];
  s.RK = in[2];
  for (int p = 0; p < 2; ++p) {
    const double *g = in + 3 + 10 * p;
    s.gm[p].r = g[0]; s.gm[p].phi = g[1]; s.gm[p].M = g[2]; s.gm[p].omega = g[3];
    s.gm[p].RK_r = g[4]; s.gm[p].RK_phi = g[5]; s.gm[p].RK_M = g[6]; s.gm[p].RK_omega = g[7];
    s.gm[p].Mdot = g[8]; s.gm[p].Macc = g[9];
  }
  e->put_scalars(d, s);
  launch_scalar(e->mesh(), e->plan, e->st, SC_MASS);
  e->check_error();
}

// ------------------------------------------------------------------ checkpoints
// src/Headers/IO.h:19-34, one rank.  The cell rows are packed and unpacked on
// the GPU (checkpoint.hip); the file layout is the reference's
// (io_hdf5.c:15-155: T, GravMass [2][4], Data [Ncells][NUM_Q + 3]).
}  // extern "C"
struct IO {
  double tcheck = 0, dtcheck = 0;
  int nfile = 0;
  char filename[256] = {0};
  double GravMassBuffer[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
  std::vector<double> buffer;
};
extern "C" {

IO *io_create(Sim *) { return new IO; }
void io_destroy(IO *io) { delete io; }
void io_allocbuf(IO *io, Sim *sim) { io->buffer.assign((size_t)sim->Ncells * (sim->NUM_Q + 3), 0.0); }
void io_deallocbuf(IO *io) { std::vector<double>().swap(io->buffer); }
double io_tcheck(IO *io) { return io->tcheck; }
int io_nfile(IO *io) { return io->nfile; }

// io_setup.c:10-19
void io_setup(IO *io, Sim *sim, TimeStep *ts) {
  io->nfile = 0;
  io->dtcheck = sim_get_T_MAX(sim) / sim->NUM_CHECKPOINTS;
  io->tcheck = io->dtcheck;
  while (io->tcheck < timestep_get_t(ts)) {
    io->tcheck += io->dtcheck;
    ++io->nfile;
  }
  snprintf(io->filename, sizeof io->filename, "checkpoint_%04d.h5", io->nfile);
}

// io_buffer.c:26-55
void io_setbuf(IO *io, Cell ***c, Sim *sim, GravMass *gm) {
  CellsHandle *h = H(c);
  mirror_push(h);
  if (io->buffer.size() != (size_t)sim->Ncells * (sim->NUM_Q + 3)) io_allocbuf(io, sim);
  if (h->eng->checkpoint_pack(0, io->buffer.data())) return;
  for (int p = 0; p < 2; ++p) {
    io->GravMassBuffer[p][0] = gm[p].r;
    io->GravMassBuffer[p][1] = gm[p].phi;
    io->GravMassBuffer[p][2] = gm[p].M;
    io->GravMassBuffer[p][3] = gm[p].omega;
  }
}

// io_buffer.c:57-81
void io_readbuf(IO *io, Cell ***c, Sim *sim, GravMass *gm) {
  CellsHandle *h = H(c);
  mirror_push(h);
  if (io->buffer.size() != (size_t)sim->Ncells * (sim->NUM_Q + 3)) {
    disco_set_error(15, "io_readbuf: no checkpoint rows (io_allocbuf, io_hdf5_in)");
    return;
  }
  if (h->eng->checkpoint_unpack(0, io->buffer.data())) return;
  h->stale = true;
  for (int p = 0; p < 2; ++p)
    gravMass_set_chkpt(gm, p, io->GravMassBuffer[p][0], io->GravMassBuffer[p][1], io->GravMassBuffer[p][2],
                       io->GravMassBuffer[p][3]);
}

// io_hdf5.c:15-155: the file named by io_setup / the previous call
void io_hdf5_out(IO *io, Sim *sim, TimeStep *ts) {
  if (!h5_available()) {
    disco_set_error(15, "checkpoints need the HDF5 library (DISCO_HDF5_LIB or libhdf5.so)");
    return;
  }
  double T = timestep_get_t(ts);
  const H5Dataset ds[3] = {{"T", 1, {1, 1}, &T},
                           {"GravMass", 2, {2, 4}, &io->GravMassBuffer[0][0]},
                           {"Data", 2, {(unsigned long long)sim->Ncells, (unsigned long long)(sim->NUM_Q + 3)},
                            io->buffer.data()}};
  if (io->buffer.size() != (size_t)sim->Ncells * (sim->NUM_Q + 3) || !h5_write_file(io->filename, ds, 3)) {
    disco_set_error(15, std::string("io_hdf5_out: cannot write ") + io->filename);
    return;
  }
  io->tcheck += io->dtcheck;
  ++io->nfile;
  snprintf(io->filename, sizeof io->filename, "checkpoint_%04d.h5", io->nfile);
}

// io_hdf5.c:158-265: input.h5 in the working directory
void io_hdf5_in(IO *io, Sim *sim, TimeStep *ts) {
  if (!h5_available()) {
    disco_set_error(15, "checkpoints need the HDF5 library (DISCO_HDF5_LIB or libhdf5.so)");
    return;
  }
  long long dims[2];
  double T = 0.0;
  const long long n = (long long)sim->Ncells * (sim->NUM_Q + 3);
  io->buffer.assign((size_t)n, 0.0);
  if (h5_read_dataset("input.h5", "T", &T, 1, dims) != 1 ||
      h5_read_dataset("input.h5", "GravMass", &io->GravMassBuffer[0][0], 8, dims) != 8 ||
      h5_read_dataset("input.h5", "Data", io->buffer.data(), n, dims) != n) {
    disco_set_error(15, "io_hdf5_in: input.h5 does not hold this mesh's T, GravMass and Data");
    return;
  }
  timestep_set_t(ts, T);
}

// launch schedules (no reference counterpart)
int disco_batch_autotune(DiscoBatch *b, int nsteps, double *ms) {
  double tmp[DISCO_NSCHEDULES];
  return b->e->autotune(nsteps, ms ? ms : tmp, DISCO_NSCHEDULES);
}
int disco_batch_set_schedule(DiscoBatch *b, int k) {
  if (k < 0 || k >= DISCO_NSCHEDULES) {
    disco_set_error(9, "schedule: 0 to DISCO_NSCHEDULES - 1");
    return -1;
  }
  b->e->set_schedule(Engine::Schedule::of(k));
  return 0;
}
int disco_batch_schedule(DiscoBatch *b) { return b->e->sched.code(); }

// the InitialDataType plugin over host arrays of one disk (host only)
int disco_initial_prim(Sim *s, const double *tiph, const double *dphi, double *prim) {
  long long NL = 0;
  for (int n : s->N_p) NL += n;
  const long long nz = NL * (long long)(s->z_faces.size() - 1);
  return disco_cell_init(s, s->InitialDataType, nullptr, tiph, dphi, prim, nz, s->NUM_Q) ? 0 : -1;
}

// the disk's InitialDataType plugin (src/main.c:103-104) over its current
// cells, on the host
int disco_batch_initial_data(DiscoBatch *b, int d) {
  Engine *e = b->e;
  if (d < 0 || d >= e->ndisk) {
    disco_set_error(9, "initial data: no such disk");
    return -1;
  }
  const long long nz = e->nzones[d];
  std::vector<double> prim((size_t)e->NQ * nz), tiph(nz), dphi(nz);
  e->download(d, 0, prim.data());
  e->download(d, 5, tiph.data());
  e->download(d, 7, dphi.data());
  Sim *S = e->sims[d];
  if (!disco_cell_init(S, S->InitialDataType, nullptr, tiph.data(), dphi.data(), prim.data(), nz, e->NQ)) return -1;
  e->upload(d, 0, prim.data());
  return 0;
}
int disco_batch_checkpoint_write(DiscoBatch *b, int d, const char *fname) { return b->e->checkpoint_write(d, fname); }
int disco_batch_checkpoint_read(DiscoBatch *b, int d, const char *fname) { return b->e->checkpoint_read(d, fname); }
}

// ------------------------------------------------------------------ unit level
// The pieces timestep_substep calls (src/TimeStep/timestep_routines.c:16-113),
// exported one by one with the reference's signatures (src/Headers/Cell.h,
// Riemann.h).  Each is a device launch on the handle's current state and on
// the unit-level conserved state (the reference's cell->cons), which follows
// the primitives until one of these calls changes it (cell_calc_cons makes
// it prim2cons of the primitives again, as do timestep_rk2 / _substep /
// _forward_euler, which end with cell_calc_cons).  The engine's own steps
// start from prim2cons of the primitives.
struct Riemann {  // riemann_setup_* fill the face, riemann_AddFlux solves it
  CellsHandle *h = nullptr;
  UnitFace f{};
  bool ready = false;
};

static int ann_of_zone(const Engine *e, long long z) {
  // annuli are in zone order: the last one starting at or before z
  int lo = 0, hi = e->nann - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) / 2;
    if (e->hann[mid].zbase <= z)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// the launch that the unit-level source routines share; part as UNIT_SRC
static void unit_src(Cell ***c, GravMass *gm, double dt, int part) {
  CellsHandle *h = H(c);
  mirror_push(h);
  Engine *e = h->eng;
  if (gm) scalars_from_host(e, nullptr, gm);
  double *uc = e->unit_cons();
  if (!uc) return;
  launch_unit(e->mesh(), e->plan, e->st, e->NQ, UNIT_SRC, uc, dt, 0.0, e->d_cool, e->d_umdot, part);
  if (part == 0 && gm) {  // gravMass_set_Mdot with the accretion summed in tile order (cell_source.c:206-213)
    std::vector<double> md(2 * (size_t)e->plan.n_ztiles);
    hip_check(hipMemcpyAsync(md.data(), e->d_umdot, md.size() * sizeof(double), hipMemcpyDeviceToHost, e->st),
              "sink D2H");
    hip_check(hipStreamSynchronize(e->st), "synchronize");
    double s0 = 0.0, s1 = 0.0;
    for (int t = 0; t < e->plan.n_ztiles; ++t) {
      s0 += md[2 * t];
      s1 += md[2 * t + 1];
    }
    gm[0].Mdot = s0;
    gm[1].Mdot = s1;
  }
}

// a plugin's single-cell function other than the disk's own cannot run on
// the device (its rows are the disk's)
static bool own_plugin(Sim *s, void (*fn)(Cell *, Sim *, GravMass *, int, int, int), const char *who) {
  if (fn != cell_single_init_ptr(s)) {
    disco_set_error(17, std::string(who) + ": only the Sim's own single-cell initial data (cell_single_init_ptr) "
                                           "runs on the device");
    return false;
  }
  return true;
}
static void unit_bc(Cell ***c, int mode, double dt) {
  CellsHandle *h = H(c);
  mirror_push(h);
  Engine *e = h->eng;
  InitRow *rows = e->unit_rows();
  if (!rows) return;
  launch_unit_bc(e->mesh(), e->plan, e->st, e->NQ, rows, mode, dt);
  h->stale = true;
}

extern "C" {

// src/Cell/cell_create_destroy.c:12-21: a cell of its own (its primitives,
// edge, width and mesh velocity; zero until set)
Cell *cell_single_create(Sim *s) {
  CellRef *r = new CellRef();
  r->h = nullptr;
  r->zone = -1;
  r->NQ = s->NUM_Q;
  r->own = new double[r->NQ + 3]();
  return reinterpret_cast<Cell *>(r);
}
void cell_single_destroy(Cell *c) {
  CellRef *r = reinterpret_cast<CellRef *>(c);
  if (!r || !r->own) return;
  delete[] r->own;
  delete r;
}

// src/Cell/cell_access.c:32-34: the cooling rate cell_add_src stored (0
// until a unit-level cell_add_src with Cooling 1 ran)
double cell_Cool(Cell ***c, int i, int j, int k) {
  CellsHandle *h = H(c);
  Engine *e = h->eng;
  if (!e->d_cool) return 0.0;
  double v = 0.0;
  hip_check(hipStreamSynchronize(e->st), "synchronize");
  hip_check(hipMemcpy(&v, e->d_cool + zone_of(h, i, j, k), sizeof(double), hipMemcpyDeviceToHost), "Cool D2H");
  return v;
}

// src/Cell/cell_modify_data.c:11-13
void cell_add_cons(Cell *cell, int q, double add) {
  CellRef *r = reinterpret_cast<CellRef *>(cell);
  if (r->own || q < 0 || q >= r->h->NQ) {
    disco_set_error(17, "cell_add_cons: a cell of a Cell *** handle and 0 <= q < NUM_Q");
    return;
  }
  mirror_push(r->h);
  Engine *e = r->h->eng;
  double *uc = e->unit_cons();
  if (!uc) return;
  launch_unit_add(uc, (long long)q * e->Ntot + r->zone, add, e->st);
}

void cell_copy(Cell ***c, Sim *) {  // src/Cell/cell_misc.c:28-41
  CellsHandle *h = H(c);
  mirror_push(h);
  Engine *e = h->eng;
  double *uc = e->unit_cons();
  if (!uc) return;
  launch_unit(e->mesh(), e->plan, e->st, e->NQ, UNIT_COPY, uc, 0.0, 0.0, nullptr, nullptr, 0);
}
void cell_adjust_RK_cons(Cell ***c, Sim *, double RK) {  // src/Cell/cell_misc.c:43-57
  CellsHandle *h = H(c);
  mirror_push(h);
  Engine *e = h->eng;
  double *uc = e->unit_cons();
  if (!uc) return;
  launch_unit(e->mesh(), e->plan, e->st, e->NQ, UNIT_ADJUST, uc, RK, 0.0, nullptr, nullptr, 0);
}

// src/Cell/cell_source.c:73-300, src/Cell/cell_add_split_fictitious.c:11-44
void cell_add_src(Cell ***c, Sim *, GravMass *gm, double dt) { unit_src(c, gm, dt, 0); }
void cell_add_visc_src(Cell ***c, Sim *, GravMass *gm, double dt) { unit_src(c, gm, dt, 1); }
void cell_add_visc_src_old(Cell ***c, Sim *, double dt) { unit_src(c, nullptr, dt, 2); }
void cell_add_split_fictitious(Cell ***c, Sim *, GravMass *gm, double dt) { unit_src(c, gm, dt, 3); }

// src/Cell/cell_misc.c:59-96
void cell_update_phi(Cell ***c, Sim *, GravMass *gm, double RK, double dt) {
  CellsHandle *h = H(c);
  mirror_push(h);
  Engine *e = h->eng;
  if (gm) scalars_from_host(e, nullptr, gm);  // the frame rotation at the masses' separation
  launch_unit(e->mesh(), e->plan, e->st, e->NQ, UNIT_UPDATE_PHI, nullptr, RK, dt, nullptr, nullptr, 0);
  h->stale = true;
}
void cell_update_dphi(Cell ***c, Sim *) {
  CellsHandle *h = H(c);
  mirror_push(h);
  Engine *e = h->eng;
  launch_unit(e->mesh(), e->plan, e->st, e->NQ, UNIT_UPDATE_DPHI, nullptr, 0.0, 0.0, nullptr, nullptr, 0);
  h->stale = true;
}

// src/Cell/cell_boundary.c:14-139 with the r faces of the last face_create
// (the ghost-annulus tile lists exist for BoundTypeR outflow)
void cell_boundary_outflow_r(Cell ***c, Face *, Sim *s, MPIsetup *, TimeStep *) {
  CellsHandle *h = H(c);
  if (s->BoundTypeR != BOUND_OUTFLOW) {
    disco_set_error(17, "cell_boundary_outflow_r: the Sim's BoundTypeR is not outflow");
    return;
  }
  mirror_push(h);
  Engine *e = h->eng;
  launch_bc_outflow_r(e->mesh(), e->plan, e->st, e->NQ);
  h->stale = true;
}
// src/Cell/cell_boundary.c:141-206 with the z faces of the last face_create
void cell_boundary_outflow_z(Cell ***c, Face *, Sim *s, MPIsetup *, TimeStep *) {
  CellsHandle *h = H(c);
  if (s->N_global[1] <= 1) return;
  mirror_push(h);
  Engine *e = h->eng;
  launch_unit(e->mesh(), e->plan, e->st, e->NQ, UNIT_OUTFLOW_Z, nullptr, 0.0, 0.0, nullptr, nullptr, 0);
  h->stale = true;
}
// src/Cell/cell_boundary.c:208-254 and src/Cell/cell_misc.c:98-132
void cell_boundary_fixed_r(Cell ***c, Sim *s, GravMass *, MPIsetup *,
                           void (*fn)(Cell *, Sim *, GravMass *, int, int, int)) {
  if (own_plugin(s, fn, "cell_boundary_fixed_r")) unit_bc(c, UNIT_FIXED_R, 0.0);
}
void cell_boundary_fixed_z(Cell ***c, Sim *s, GravMass *, MPIsetup *,
                           void (*fn)(Cell *, Sim *, GravMass *, int, int, int)) {
  if (own_plugin(s, fn, "cell_boundary_fixed_z")) unit_bc(c, UNIT_FIXED_Z, 0.0);
}
void cell_bc_damp(Cell ***c, Sim *s, GravMass *, double dt, void (*fn)(Cell *, Sim *, GravMass *, int, int, int)) {
  if (!(s->DAMP_TIME > 0.0)) {
    disco_set_error(17, "cell_bc_damp: DAMP_TIME must be positive (the damping rate is 1/DAMP_TIME/2/pi)");
    return;
  }
  if (own_plugin(s, fn, "cell_bc_damp")) unit_bc(c, UNIT_DAMP, dt);
}

// declared at src/Headers/Cell.h:91 and never defined in the reference: the
// cell's primitives, edge and width on stdout
void cell_print(Cell ***c, int i, int j, int k) {
  Cell *x = cell_single(c, i, j, k);
  printf("cell (%d %d %d): tiph %e dphi %e prim", i, j, k, cell_tiph(x), cell_dphi(x));
  for (int q = 0; q < H(c)->NQ; ++q) printf(" %e", cell_prim(x, q));
  printf("\n");
}

// ------------------------------------------------------------------ Riemann
// src/Riemann/Riemann_create_destroy.c:9-53, Riemann_setup.c:11-80,
// Riemann_routines.c:314-397: one face at a time, solved on the device
Riemann *riemann_create(Sim *) { return new Riemann(); }
void riemann_destroy(Riemann *r) { delete r; }
void riemann_setup_rz(Riemann *R, Face *F, Sim *, int n, int direction) {
  if (!F || n < 0 || n >= (int)F->L.size() || (direction != 0 && direction != 2)) {
    disco_set_error(18, "riemann_setup_rz: no such face (direction 0 for r faces, 2 for z)");
    R->ready = false;
    return;
  }
  R->h = F->h;
  UnitFace &f = R->f;
  f.kind = direction;
  f.cL = F->L[n];
  f.cR = F->R[n];
  f.ann_L = ann_of_zone(F->h->eng, f.cL);
  f.r = F->r[n];
  f.dA = F->dA[n];
  f.cm = F->cm[n];
  f.deltaL = F->deltaL[n];
  f.deltaR = F->deltaR[n];
  f.r_cell_L = direction == 0 ? f.r - f.deltaL : f.r;
  f.r_cell_R = direction == 0 ? f.r + f.deltaR : f.r;
  R->ready = true;
}
void riemann_setup_p(Riemann *R, Cell ***c, Sim *s, int i, int j_low, int k, int) {
  CellsHandle *h = H(c);
  const int N = sim_N_p(s, i);
  if (i < 0 || i >= sim_N(s, 0) || k < 0 || k >= sim_N(s, 1) || j_low < 0 || j_low >= N) {
    disco_set_error(18, "riemann_setup_p: no such cell");
    R->ready = false;
    return;
  }
  const int j_hi = j_low == N - 1 ? 0 : j_low + 1;
  R->h = h;
  UnitFace &f = R->f;
  f.kind = 1;
  f.cL = zone_of(h, i, j_low, k);
  f.cR = zone_of(h, i, j_hi, k);
  f.ann_L = k * h->eng->hdp[0].NR + i;
  const double zm = sim_FacePos(s, k - 1, 1), zp = sim_FacePos(s, k, 1);
  const double dz = zp - zm;
  const double rm = sim_FacePos(s, i - 1, 0), rp = sim_FacePos(s, i, 0);
  const double dr = rp - rm;
  const double r = .5 * (rp + rm);
  f.dA = dr * dz;
  f.r = r;
  f.cm = 0.0;  // the left cell's tiph, read on the device
  f.deltaL = f.deltaR = 0.0;
  f.r_cell_L = r;
  f.r_cell_R = r;
  R->ready = true;
}
void riemann_AddFlux(Riemann *R, Sim *, GravMass *gm, double dt) {
  if (!R->ready) {
    disco_set_error(18, "riemann_AddFlux before riemann_setup_rz / riemann_setup_p");
    return;
  }
  CellsHandle *h = R->h;
  mirror_push(h);
  Engine *e = h->eng;
  if (gm) scalars_from_host(e, nullptr, gm);
  double *uc = e->unit_cons();
  if (!uc) return;
  launch_unit_face(e->mesh(), e->st, e->NQ, uc, R->f, dt);
}

// declared at src/Headers/Diagnostics.h:31 and never defined in the
// reference: starts new time averages at the current time
void diagnostics_reset(Diagnostics *d, Cell ***, Sim *, TimeStep *ts) {
  std::fill(d->vec.begin(), d->vec.end(), 0.0);
  std::fill(d->scal.begin(), d->scal.end(), 0.0);
  d->toutprev = d->toutprev_dump = timestep_get_t(ts);
}
// declared at src/Headers/Sim.h:89 and never defined in the reference; no
// parfile key sets it (src/Sim/sim_readpar.c:55-106): 0
int sim_ZeroPsiBndry(Sim *) { return 0; }

}  // extern "C"

extern "C" {
// the conservation ledger of a class-API handle (no reference counterpart)
int disco_cell_ledger_enable(Cell ***c, int on) { return H(c)->eng->ledger_enable(on != 0); }
int disco_cell_ledger(Cell ***c, double *out) { return H(c)->eng->ledger_last(0, out); }
}
