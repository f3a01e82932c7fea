// Host-side classes: MPIsetup (one rank), Sim (parameters and mesh
// geometry) and GravMass (the two point masses).  The mesh must be bitwise
// identical to the reference's, so the arithmetic below keeps the
// reference's expression order; it runs on the host libm like the reference.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>

#include "../../include/disco.h"
#include "disco_internal.h"

#define TWO_PI (2.0 * M_PI)  // the reference's PHIMAX, src/Headers/header.h:21

// ------------------------------------------------------------------ MPIsetup
// src/MPIsetup/mpisetup_*.c on a single rank: the Cartesian communicator is
// 1x1 with wraparound, so every rank-boundary predicate is true.
extern "C" {
MPIsetup *mpisetup_create(int, char **) { return new MPIsetup(); }
void mpisetup_destroy(MPIsetup *m) { delete m; }
void mpisetup_setprocs(MPIsetup *, char *) {}
void mpisetup_cart_create(MPIsetup *) {}
void mpisetup_left_right(MPIsetup *) {}
int mpisetup_check_rin_bndry(MPIsetup *m) { return m->dim_MyProc[0] == 0; }
int mpisetup_check_rout_bndry(MPIsetup *m) { return m->dim_MyProc[0] == m->dim_NumProcs[0] - 1; }
int mpisetup_check_zbot_bndry(MPIsetup *m) { return m->dim_MyProc[1] == 0; }
int mpisetup_check_ztop_bndry(MPIsetup *m) { return m->dim_MyProc[1] == m->dim_NumProcs[1] - 1; }
int mpisetup_MyProc(MPIsetup *m) { return m->MyProc; }
int mpisetup_NumProcs(MPIsetup *m) { return m->NumProcs; }
int mpisetup_dim_MyProc(MPIsetup *m, int d) { return m->dim_MyProc[d]; }
int mpisetup_dim_NumProcs(MPIsetup *m, int d) { return m->dim_NumProcs[d]; }
int mpisetup_left_Proc(MPIsetup *m, int d) { return m->left_Proc[d]; }
int mpisetup_right_Proc(MPIsetup *m, int d) { return m->right_Proc[d]; }
}

// ------------------------------------------------------------------ parameters
namespace {
enum PType { PI, PD };
struct PKey {
  const char *key;
  PType type;
  size_t off;
  bool required;
};
#define PK_I(k, f) {k, PI, offsetof(Sim, f), true}
#define PK_D(k, f) {k, PD, offsetof(Sim, f), true}
// key names and order of src/Sim/sim_readpar.c:55-106
const PKey kKeys[] = {
    PK_I("Restart", Restart), PK_I("InitialDataType", InitialDataType),
    PK_I("GravMassType", GravMassType), PK_D("MassRatio", MassRatio),
    PK_D("OrbShrinkTscale", OrbShrinkTscale), PK_D("OrbShrinkT0", OrbShrinkT0),
    PK_I("BoundTypeR", BoundTypeR), PK_I("BoundTypeZ", BoundTypeZ), PK_I("NoInnerBC", NoInnerBC),
    PK_I("NumR", N_global[0]), PK_I("NumZ", N_global[1]), PK_I("ng", ng),
    PK_D("R_Min", MIN[0]), PK_D("R_Max", MAX[0]), PK_D("Z_Min", MIN[1]), PK_D("Z_Max", MAX[1]),
    PK_I("NP_CONST", NP_CONST), PK_I("NPCAP", NPCAP), PK_D("aspect", aspect),
    PK_I("NUM_C", NUM_C), PK_I("NUM_N", NUM_N), PK_D("Time_Max", T_MAX),
    PK_I("Num_Checkpoints", NUM_CHECKPOINTS), PK_I("Num_Diag_Dump", NUM_DIAG_DUMP),
    PK_I("Num_Diag_Measure", NUM_DIAG_MEASURE), PK_I("w_numeric_type", W_NUMERIC_TYPE),
    PK_I("RiemannSolver", Riemann), PK_D("Adiabatic_Index", GAMMALAW), PK_I("Cooling", COOLING),
    PK_D("Explicit_Viscosity", EXPLICIT_VISCOSITY), PK_I("ViscConst", VISC_CONST),
    PK_D("PoRho_r1", PoRho_r1), PK_D("CFL", CFL), PK_D("PLM", PLM), PK_I("Grav_2D", GRAV2D),
    PK_D("G_EPS", G_EPS), PK_I("RhoSinkOn", RhoSinkOn), PK_I("PHI_ORDER", PHI_ORDER),
    PK_D("Rho_Floor", RHO_FLOOR), PK_D("Cs_Floor", CS_FLOOR), PK_D("Cs_Cap", CS_CAP),
    PK_D("Vel_Cap", VEL_CAP), PK_D("DAMP_TIME", DAMP_TIME), PK_D("RDAMP_INNER", RDAMP_INNER),
    PK_D("RDAMP_OUTER", RDAMP_OUTER), PK_D("RLogScale", RLogScale), PK_D("ZLogScale", ZLogScale),
    PK_D("HiResSigma", HiResSigma), PK_D("HiResR0a", HiResR0a), PK_D("HiResR0b", HiResR0b),
    PK_D("HiResFac", HiResFac), PK_I("w_analytic_type", W_ANALYTIC_TYPE),
    {"N_p_round", PI, offsetof(Sim, N_p_round), false},
};

void put_value(Sim *s, const PKey &k, double v) {
  char *base = reinterpret_cast<char *>(s) + k.off;
  if (k.type == PI)
    *reinterpret_cast<int *>(base) = (int)v;
  else
    *reinterpret_cast<double *>(base) = v;
}

// readvar, src/Sim/sim_readpar.c:10-46: first line whose first token is
// the key; the value follows the key after any of "\t :=>_".
int read_key(const char *filename, const char *varname, double *out) {
  FILE *in = fopen(filename, "r");
  if (!in) return 2;
  char s[512], nm[512], s1[512];
  int found = 0;
  while (!found && fgets(s, 512, in) != NULL) {
    nm[0] = 0;
    if (sscanf(s, "%511s ", nm) != 1) continue;
    if (strcmp(nm, varname) == 0) {
      strcpy(s1, s);
      found = 1;
    }
  }
  fclose(in);
  if (!found) return 1;
  char *s2 = s1 + strlen(nm) + strspn(s1 + strlen(nm), "\t :=>_");
  double temp = 0.0;
  if (sscanf(s2, "%lf", &temp) != 1) return 1;
  *out = temp;
  return 0;
}
}  // namespace

// ------------------------------------------------------------------ Sim
extern "C" {
Sim *sim_create(MPIsetup *) { return new Sim(); }

void sim_destroy(Sim *s) { delete s; }

int sim_read_par_file(Sim *s, MPIsetup *, char *filename) {
  int err = 0;
  std::string missing;
  for (const PKey &k : kKeys) {
    double v;
    int e = read_key(filename, k.key, &v);
    if (e == 2) {
      disco_set_error(3, std::string("cannot open parameter file ") + filename);
      return 1;
    }
    if (e == 0)
      put_value(s, k, v);
    else if (k.required) {
      ++err;
      missing += std::string(" ") + k.key;
    }
  }
  if (err) {
    disco_set_error(3, "parameter file lacks:" + missing);
    return 1;
  }
  return 0;
}

int sim_set_param(Sim *s, const char *key, double value) {
  for (const PKey &k : kKeys)
    if (strcmp(k.key, key) == 0) {
      put_value(s, k, value);
      return 0;
    }
  return 1;
}

// src/Sim/sim_alloc_arr.c:9-46 on the rank that holds both global edges
void sim_alloc_arr(Sim *s, MPIsetup *m) {
  s->N_noghost[0] = s->N_global[0] / m->dim_NumProcs[0];
  s->N_noghost[1] = s->N_global[1] / m->dim_NumProcs[1];
  s->Nghost_min[0] = (m->dim_MyProc[0] == 0) ? 1 : s->ng;
  s->Nghost_max[0] = s->ng;
  if (s->N_global[1] == 1) {
    s->Nghost_min[1] = 0;
    s->Nghost_max[1] = 0;
  } else {
    s->Nghost_min[1] = s->ng;
    s->Nghost_max[1] = s->ng;
  }
  int NR = s->N_noghost[0] + s->Nghost_min[0] + s->Nghost_max[0];
  int NZ = s->N_noghost[1] + s->Nghost_min[1] + s->Nghost_max[1];
  s->N_p.assign(NR, 0);
  s->r_faces.assign(NR + 1, 0.0);
  s->z_faces.assign(NZ + 1, 0.0);
  s->allocated = true;
}

double sim_MIN(Sim *s, int d) { return s->MIN[d]; }
double sim_MAX(Sim *s, int d) { return s->MAX[d]; }
int sim_N0(Sim *s, int d) { return s->N0[d]; }
int sim_N_p(Sim *s, int i) { return s->N_p[i]; }
double sim_FacePos(Sim *s, int index, int direction) {  // src/Sim/sim_access.c:26-36
  if (direction == 0) return s->r_faces[index + 1];
  if (direction == 1) return s->z_faces[index + 1];
  disco_set_error(4, "sim_FacePos: bad direction");
  return 0.0;
}
int sim_N(Sim *s, int d) { return s->N_noghost[d] + s->Nghost_min[d] + s->Nghost_max[d]; }
int sim_NoInnerBC(Sim *s) { return s->NoInnerBC; }
int sim_Restart(Sim *s) { return s->Restart; }
int sim_BoundTypeR(Sim *s) { return s->BoundTypeR; }
int sim_BoundTypeZ(Sim *s) { return s->BoundTypeZ; }
int sim_N_global(Sim *s, int d) { return s->N_global[d]; }
int sim_Ncells(Sim *s) { return s->Ncells; }
int sim_Ncells_global(Sim *s) { return s->Ncells_global; }
int sim_offset(Sim *s) { return s->offset; }
int sim_ng(Sim *s) { return s->ng; }
int sim_Nghost_min(Sim *s, int d) { return s->Nghost_min[d]; }
int sim_Nghost_max(Sim *s, int d) { return s->Nghost_max[d]; }
int sim_W_NUMERIC_TYPE(Sim *s) { return s->W_NUMERIC_TYPE; }
int sim_NumGravMass(Sim *s) { return s->NumGravMass; }
double sim_MassRatio(Sim *s) { return s->MassRatio; }
double sim_OrbShrinkTscale(Sim *s) { return s->OrbShrinkTscale; }
double sim_OrbShrinkT0(Sim *s) { return s->OrbShrinkT0; }
int sim_Riemann(Sim *s) { return s->Riemann; }
double sim_GAMMALAW(Sim *s) { return s->GAMMALAW; }
int sim_COOLING(Sim *s) { return s->COOLING; }
double sim_EXPLICIT_VISCOSITY(Sim *s) { return s->EXPLICIT_VISCOSITY; }
int sim_VISC_CONST(Sim *s) { return s->VISC_CONST; }
double sim_PoRho_r1(Sim *s) { return s->PoRho_r1; }
double sim_CFL(Sim *s) { return s->CFL; }
double sim_PLM(Sim *s) { return s->PLM; }
int sim_W_ANALYTIC_TYPE(Sim *s) { return s->W_ANALYTIC_TYPE; }
int sim_GRAV2D(Sim *s) { return s->GRAV2D; }
double sim_G_EPS(Sim *s) { return s->G_EPS; }
int sim_RhoSinkOn(Sim *s) { return s->RhoSinkOn; }
int sim_PHI_ORDER(Sim *s) { return s->PHI_ORDER; }
double sim_RHO_FLOOR(Sim *s) { return s->RHO_FLOOR; }
double sim_CS_FLOOR(Sim *s) { return s->CS_FLOOR; }
double sim_CS_CAP(Sim *s) { return s->CS_CAP; }
double sim_VEL_CAP(Sim *s) { return s->VEL_CAP; }
int sim_NUM_C(Sim *s) { return s->NUM_C; }
int sim_NUM_Q(Sim *s) { return s->NUM_Q; }
double sim_get_T_MAX(Sim *s) { return s->T_MAX; }
int sim_NUM_CHECKPOINTS(Sim *s) { return s->NUM_CHECKPOINTS; }
int sim_NUM_DIAG_DUMP(Sim *s) { return s->NUM_DIAG_DUMP; }
int sim_NUM_DIAG_MEASURE(Sim *s) { return s->NUM_DIAG_MEASURE; }
int sim_InitialDataType(Sim *s) { return s->InitialDataType; }
double sim_DAMP_TIME(Sim *s) { return s->DAMP_TIME; }
double sim_RDAMP_INNER(Sim *s) { return s->RDAMP_INNER; }
double sim_RDAMP_OUTER(Sim *s) { return s->RDAMP_OUTER; }
int sim_GravMassType(Sim *s) { return s->GravMassType; }
}

// root finder of the hi-res radial map, src/Sim/sim_set.c:36-71
static double r_func(double r, double r2, double fac, double sigma, double r0a, double r0b) {
  double F = r + fac * sigma * sqrt(M_PI / 4.0) * (erf((r - r0a) / sigma) + 1.0) +
             fac * sigma * sqrt(M_PI / 4.0) * (erf((r - r0b) / sigma) + 1.0) - r2;
  return F;
}
static int sgn(double x) {
  if (x > 0) return 1;
  if (x < 0) return -1;
  return 0;
}
static double get_r(double r2, double RMIN, double RMAX, double r0a, double r0b, double sigma, double fac) {
  double r = 0.0;
  int n = 1;
  const int NMAX = 1000;
  double a = RMIN, b = RMAX;
  const double TOL = 1.0e-12;
  while (n <= NMAX) {
    double c = (a + b) / 2.;
    if ((r_func(c, r2, fac, sigma, r0a, r0b) == 0.0) || ((b - a) / 2. < TOL)) {
      r = c;
      break;
    }
    n = n + 1;
    if (sgn(r_func(c, r2, fac, sigma, r0a, r0b)) == sgn(r_func(a, r2, fac, sigma, r0a, r0b)))
      a = c;
    else
      b = c;
  }
  return r;
}

extern "C" {
// face positions, src/Sim/sim_set.c:73-143
void sim_set_rz(Sim *s, MPIsetup *m) {
  s->N0[0] = s->N_noghost[0] * m->dim_MyProc[0];
  s->N0[1] = s->N_noghost[1] * m->dim_MyProc[1];
  double RMIN = s->MIN[0], ZMIN = s->MIN[1], RMAX = s->MAX[0], ZMAX = s->MAX[1];
  double Rscale = s->RLogScale, sigma = s->HiResSigma, r0a = s->HiResR0a, r0b = s->HiResR0b;
  double fac = s->HiResFac;
  double r2_max = RMAX + fac * sigma * sqrt(M_PI / 4.0) * (erf((RMAX - r0a) / sigma) + 1.0) +
                  fac * sigma * sqrt(M_PI / 4.0) * (erf((RMAX - r0b) / sigma) + 1.0);
  double r2_min = RMIN + fac * sigma * sqrt(M_PI / 4.0) * (erf((RMIN - r0a) / sigma) + 1.0) +
                  fac * sigma * sqrt(M_PI / 4.0) * (erf((RMIN - r0b) / sigma) + 1.0);
  double r1_max = Rscale * log(1.0 + r2_max / Rscale);
  double r1_min = Rscale * log(1.0 + r2_min / Rscale);
  int NR = sim_N(s, 0), NZ = sim_N(s, 1);
  for (int i = 0; i < NR + 1; i++) {
    int ig = i - s->Nghost_min[0] + s->N0[0];
    double delta = (r1_max - r1_min) / (double)s->N_global[0];
    double r1 = r1_min + (double)ig * delta;
    double r2 = Rscale * (exp(r1 / Rscale) - 1.0);
    s->r_faces[i] = get_r(r2, 0., 2. * RMAX, r0a, r0b, sigma, fac);
  }
  if (s->NoInnerBC == 1 && mpisetup_check_rin_bndry(m)) s->r_faces[0] = 0.0;
  double Zscale = s->ZLogScale;
  double z1_max = sgn(ZMAX) * Zscale * log(1.0 + fabs(ZMAX) / Zscale);
  double z1_min = sgn(ZMIN) * Zscale * log(1.0 + fabs(ZMIN) / Zscale);
  for (int k = 0; k < NZ + 1; k++) {
    int kg = k - s->Nghost_min[1] + s->N0[1];
    double delta = (z1_max - z1_min) / (double)s->N_global[1];
    double z1 = z1_min + (double)kg * delta;
    s->z_faces[k] = sgn(z1) * Zscale * (exp(fabs(z1) / Zscale) - 1.0);
  }
}

// src/Sim/sim_set.c:10-33; with the extra key N_p_round = 1 the count is
// floor(2 pi r_c / dr + 0.5) instead of the truncated aspect formula.
void sim_set_N_p(Sim *s) {
  int NR = sim_N(s, 0);
  if (s->NP_CONST > 0) {
    for (int i = 0; i < NR; ++i) s->N_p[i] = s->NP_CONST;
    return;
  }
  for (int i = 0; i < NR; ++i) {
    double r_cell = 0.5 * (sim_FacePos(s, i - 1, 0) + sim_FacePos(s, i, 0));
    double dr = sim_FacePos(s, i, 0) - sim_FacePos(s, i - 1, 0);
    if (r_cell < s->MIN[0]) dr = sim_FacePos(s, i + 1, 0) - sim_FacePos(s, i, 0);
    if (s->N_p_round)
      s->N_p[i] = (int)floor(2.0 * M_PI * r_cell / dr + 0.5);
    else
      s->N_p[i] = (int)(TWO_PI * (1. + (r_cell / dr - 1.) / s->aspect));
    if (s->NPCAP > 0 && s->N_p[i] > s->NPCAP) s->N_p[i] = s->NPCAP;
  }
}

void sim_set_N_p_array(Sim *s, const int *N_p) {
  for (size_t i = 0; i < s->N_p.size(); ++i) s->N_p[i] = N_p[i];
}

// src/Sim/sim_set.c:145-177 on one rank
void sim_set_misc(Sim *s, MPIsetup *) {
  int NR = sim_N(s, 0), NZ = sim_N(s, 1);
  int Ncells = 0;
  for (int i = 0; i < NR; ++i) {
    if (s->N_p[i] < 2) {
      // the reference goes on with a negative count and crashes later (an
      // inner ghost face at r < 0 when R_Min is below the radial spacing)
      disco_set_error(8, "annulus " + std::to_string(i) + " has N_p = " + std::to_string(s->N_p[i]) +
                             ": every annulus needs N_p >= 2 (check that R_Min exceeds the radial spacing)");
      return;
    }
    Ncells += s->N_p[i] * NZ;
  }
  s->Ncells = Ncells;
  s->Ncells_global = Ncells;
  s->offset = 0;
  s->NumGravMass = 2;
  s->NUM_Q = s->NUM_C + s->NUM_N;
}
}

// analytic frame rotation: disco_rOm_a and friends (frame.cpp)
extern "C" {
double sim_rOm_a(Sim *s, double r, double a) { return disco_rOm_a(s->W_ANALYTIC_TYPE, r, a); }
double sim_rdrOm_a(Sim *s, double r, double a) { return disco_rdrOm_a(s->W_ANALYTIC_TYPE, r, a); }
double sim_dtOm_a(Sim *s, double r, double a) {
  return disco_dtOm_a(s->W_ANALYTIC_TYPE, r, a, s->OrbShrinkTscale);
}
}

// ------------------------------------------------------------------ phi offsets
// The reference seeds the C library generator with srand(666 + rank) and
// draws one offset per (k, i) annulus, k outer, i inner
// (src/Cell/cell_create_destroy.c:64-71); GlibcRand restates that generator.

// tiph and dphi of every zone exactly as cell_create leaves them
// (src/Cell/cell_create_destroy.c:64-91) followed by cell_clean_pi
// (src/Cell/cell_misc.c:12-26) and, on one rank, the z-periodic sync of
// tiph (src/Cell/cell_sync.c:156-205) done by main (src/main.c:80-85).
void disco_cell_offsets(const Sim *s, std::vector<double> &tiph, std::vector<double> &dphi) {
  int NR = (int)s->N_p.size();
  int NZ = (int)s->z_faces.size() - 1;
  long long NL = 0;
  for (int i = 0; i < NR; ++i) {
    if (s->N_p[i] < 1) {
      disco_set_error(8, "cell offsets: annulus with N_p < 1");
      tiph.clear();
      dphi.clear();
      return;
    }
    NL += s->N_p[i];
  }
  tiph.assign((size_t)NL * NZ, 0.0);
  dphi.assign((size_t)NL * NZ, 0.0);
  GlibcRand g(666);
  const double RAND_MAX_D = 2147483647.0;
  long long z = 0;
  for (int k = 0; k < NZ; ++k)
    for (int i = 0; i < NR; ++i) {
      double tiph_0 = TWO_PI * (double)g.next() / RAND_MAX_D;
      double d = TWO_PI / (double)s->N_p[i];
      for (int j = 0; j < s->N_p[i]; ++j, ++z) {
        double t = tiph_0 + (double)j * d;
        while (t > TWO_PI) t -= TWO_PI;
        while (t < 0.0) t += TWO_PI;
        tiph[z] = t;
        dphi[z] = d;
      }
    }
  if (s->N_global[1] > 1) {
    int ng_max = s->Nghost_max[1], ng_min = s->Nghost_min[1];
    int kNm2g = NZ - 2 * ng_max, kNmg = NZ - ng_max, kg = ng_min;
    for (int d = 0; d < ng_max; ++d)
      for (long long x = 0; x < NL; ++x) tiph[(kNmg + d) * NL + x] = tiph[(kg + d) * NL + x];
    for (int d = 0; d < ng_min; ++d)
      for (long long x = 0; x < NL; ++x) tiph[d * NL + x] = tiph[(kNm2g + d) * NL + x];
  }
}

extern "C" void disco_initial_edges(Sim *s, double *tiph, double *dphi) {
  std::vector<double> t, d;
  disco_cell_offsets(s, t, d);
  memcpy(tiph, t.data(), t.size() * sizeof(double));
  memcpy(dphi, d.data(), d.size() * sizeof(double));
}

void sim_param_arrays(const Sim *s, DiskParams *p) {
  p->NR = (int)s->N_p.size();
  p->NZ = (int)s->z_faces.size() - 1;
  p->NQ = s->NUM_Q;
  p->NC = s->NUM_C;
  p->NgMinR = s->Nghost_min[0];
  p->NgMaxR = s->Nghost_max[0];
  p->NgMinZ = s->Nghost_min[1];
  p->NgMaxZ = s->Nghost_max[1];
  p->NglobalZ = s->N_global[1];
  p->w_numeric = s->W_NUMERIC_TYPE;
  p->w_analytic = s->W_ANALYTIC_TYPE;
  p->riemann = s->Riemann;
  p->visc_const = s->VISC_CONST;
  p->grav2d = s->GRAV2D;
  p->sink = s->RhoSinkOn;
  p->phi_order = s->PHI_ORDER;
  p->cooling = s->COOLING;
  p->initial_data_type = s->InitialDataType;
  p->grav_type = s->GravMassType;
  p->bound_r = s->BoundTypeR;
  p->no_inner_bc = s->NoInnerBC;
  p->gamma = s->GAMMALAW;
  p->cfl = s->CFL;
  p->plm = s->PLM;
  p->visc = s->EXPLICIT_VISCOSITY;
  p->g_eps = s->G_EPS;
  p->rho_floor = s->RHO_FLOOR;
  p->cs_floor = s->CS_FLOOR;
  p->cs_cap = s->CS_CAP;
  p->vel_cap = s->VEL_CAP;
  p->orb_tscale = s->OrbShrinkTscale;
  p->orb_t0 = s->OrbShrinkT0;
  p->t_max = s->T_MAX;
  p->porho_r1 = s->PoRho_r1;
}

// ------------------------------------------------------------------ GravMass
// src/GravMass/GravMass_*.c; always two masses (src/Sim/sim_set.c:173)
extern "C" {
GravMass *gravMass_create(int n) {
  GravMass *g = new GravMass[n < 2 ? 2 : n];
  memset(g, 0, sizeof(GravMass) * (n < 2 ? 2 : n));
  return g;
}
void gravMass_destroy(GravMass *g) { delete[] g; }

void gravMass_init_none(GravMass *g, Sim *) {
  for (int p = 0; p < 2; ++p) {
    g[p].M = 0.0; g[p].r = 0.0; g[p].phi = 0.0; g[p].omega = 0.0; g[p].Mdot = 0.0; g[p].Macc = 0.0;
  }
}
void gravMass_init_single(GravMass *g, Sim *) {
  g[0].M = 1.0; g[0].r = 0.0; g[0].phi = 0.0; g[0].omega = 0.0; g[0].Mdot = 0.0; g[0].Macc = 0.0;
  g[1].M = 0.0; g[1].r = 1.0; g[1].phi = 0.0; g[1].omega = 0.0; g[1].Mdot = 0.0; g[1].Macc = 0.0;
}
void gravMass_init_binary(GravMass *g, Sim *s) {  // GravMass_init.c:53-84
  double Mtotal = 1.0, sep = 1.0;
  double massratio = s->MassRatio;
  double M0 = Mtotal / (1. + massratio);
  double M1 = Mtotal / (1. + 1. / massratio);
  double r0 = M1 / Mtotal * sep;
  double r1 = M0 / Mtotal * sep;
  double om2 = 1. / sep / sep / sep;
  for (int p = 0; p < 2; ++p) {
    g[p].OrbShrinkTscale = s->OrbShrinkTscale;
    g[p].OrbShrinkT0 = s->OrbShrinkT0;
    g[p].omega = sqrt(om2);
    g[p].Mdot = 0.0;
    g[p].Macc = 0.0;
  }
  g[0].M = M0; g[0].r = r0; g[0].phi = 0.0;
  g[1].M = M1; g[1].r = r1; g[1].phi = M_PI;
}
void (*gravMass_init_ptr(Sim *s))(GravMass *, Sim *) {
  if (s->GravMassType == 0) return &gravMass_init_none;
  if (s->GravMassType == 1) return &gravMass_init_single;
  if (s->GravMassType == 2) return &gravMass_init_binary;
  disco_set_error(6, "unknown GravMassType");
  return &gravMass_init_none;
}
void gravMass_set_chkpt(GravMass *g, int p, double r, double phi, double M, double omega) {
  g[p].r = r; g[p].phi = phi; g[p].M = M; g[p].omega = omega;
}
double gravMass_r(GravMass *g, int p) { return g[p].r; }
double gravMass_phi(GravMass *g, int p) { return g[p].phi; }
double gravMass_M(GravMass *g, int p) { return g[p].M; }
double gravMass_omega(GravMass *g, int p) { return g[p].omega; }
double gravMass_Mdot(GravMass *g, int p) { return g[p].Mdot; }
double gravMass_Macc(GravMass *g, int p) { return g[p].Macc; }
double gravMass_dist(GravMass *g, int p, double r, double phi, double z) {
  double xbh = g[p].r * cos(g[p].phi);
  double ybh = g[p].r * sin(g[p].phi);
  double x = r * cos(phi);
  double y = r * sin(phi);
  return sqrt((x - xbh) * (x - xbh) + (y - ybh) * (y - ybh) + z * z);
}
void gravMass_clean_pi(GravMass *g, Sim *) {
  for (int p = 0; p < 2; ++p) {
    double phi = g[p].phi;
    while (phi > 2. * M_PI) phi -= 2. * M_PI;
    while (phi < 0.0) phi += 2. * M_PI;
    g[p].phi = phi;
  }
}
void gravMass_copy(GravMass *g, Sim *) {
  for (int p = 0; p < 2; ++p) {
    g[p].RK_r = g[p].r; g[p].RK_phi = g[p].phi; g[p].RK_M = g[p].M; g[p].RK_omega = g[p].omega;
  }
}
void gravMass_move(Sim *s, GravMass *g, double t, double dt) {  // GravMass_misc.c:31-64
  double m0 = g[0].M, m1 = g[1].M;
  double M = m0 + m1;
  double a_0 = 1.0, tau_0 = s->OrbShrinkTscale, T0 = s->OrbShrinkT0;
  double a, omega;
  if ((t - T0) / tau_0 < 1.) {
    a = a_0 * pow((1.0 - (t - T0) / tau_0), 0.25);
    omega = pow(a / M, -1.5);
  } else {
    a = 0.0;
    omega = 0.0;
  }
  double r0, r1;
  if (M < 1.e-12) {
    r0 = 0.0;
    r1 = 0.0;
  } else {
    r0 = m1 / M * a;
    r1 = m0 / M * a;
  }
  g[0].phi += omega * dt;
  g[1].phi += omega * dt;
  g[0].r = r0;
  g[1].r = r1;
  g[0].omega = omega;
  g[1].omega = omega;
}
void gravMass_update_RK(GravMass *g, Sim *, double RK) {
  for (int i = 0; i < 2; ++i) {
    while (g[i].phi - g[i].RK_phi > M_PI) g[i].RK_phi += 2. * M_PI;
    while (g[i].phi - g[i].RK_phi < -M_PI) g[i].RK_phi -= 2. * M_PI;
    g[i].r = (1.0 - RK) * g[i].r + RK * g[i].RK_r;
    g[i].phi = (1.0 - RK) * g[i].phi + RK * g[i].RK_phi;
    g[i].M = (1.0 - RK) * g[i].M + RK * g[i].RK_M;
    g[i].omega = (1.0 - RK) * g[i].omega + RK * g[i].RK_omega;
  }
}
void gravMass_set_Mdot(GravMass *g, double v, int p) { g[p].Mdot = v; }
void gravMass_set_Macc(GravMass *g, double v, int p) { g[p].Macc = v; }
}
