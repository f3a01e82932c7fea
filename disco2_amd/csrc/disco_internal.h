// Internal host-side definitions of the opaque classes of include/disco.h.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "disco_types.h"

struct Engine;

struct MPIsetup {  // one rank; src/Headers/MPIsetup.h:10-20 fields
  int ndims_mpi = 1;
  int MyProc = 0, NumProcs = 1;
  int dim_MyProc[2] = {0, 0};
  int dim_NumProcs[2] = {1, 1};
  int left_Proc[2] = {0, 0};
  int right_Proc[2] = {0, 0};
};

struct Sim {  // src/Headers/Sim.h:8-71 fields
  std::vector<double> r_faces, z_faces;
  std::vector<int> N_p;
  int NPCAP = 0;
  int N0[2] = {0, 0};
  int N_noghost[2] = {0, 0};
  int Nghost_min[2] = {0, 0};
  int Nghost_max[2] = {0, 0};
  int Ncells = 0, Ncells_global = 0, offset = 0;
  int InitialDataType = 0, GravMassType = 0, BoundTypeR = 0, BoundTypeZ = 0, NoInnerBC = 0, Restart = 0;
  int N_global[2] = {0, 0};
  int NP_CONST = 0;
  double aspect = 1.0;
  int ng = 0;
  double T_MAX = 0;
  int NUM_CHECKPOINTS = 0, NUM_DIAG_DUMP = 0, NUM_DIAG_MEASURE = 0;
  double MIN[2] = {0, 0}, MAX[2] = {0, 0};
  int NUM_C = 5, NUM_N = 0, NUM_Q = 5;
  int W_NUMERIC_TYPE = 0;
  int NumGravMass = 2;
  double MassRatio = 1, OrbShrinkTscale = 1e19, OrbShrinkT0 = 0;
  int Riemann = 0;
  double GAMMALAW = 1.4;
  int COOLING = 0;
  double EXPLICIT_VISCOSITY = -1;
  int VISC_CONST = 0;
  double PoRho_r1 = 1, CFL = 0.4, PLM = 1;
  int GRAV2D = 1;
  double G_EPS = 0;
  int RhoSinkOn = 0, PHI_ORDER = 2;
  double RHO_FLOOR = 0, CS_FLOOR = 0, CS_CAP = 0, VEL_CAP = 0;
  double DAMP_TIME = -1, RDAMP_INNER = 0, RDAMP_OUTER = 0;
  double RLogScale = 1e6, ZLogScale = 1e6, HiResSigma = 1e6, HiResR0a = 1e6, HiResR0b = 1e6, HiResFac = 0;
  int W_ANALYTIC_TYPE = 0;
  int N_p_round = 0;  // extra key: N_phi = floor(2 pi r_c / dr + 0.5)
  bool allocated = false;
};

struct GravMass {  // src/Headers/GravMass.h:10-22 fields
  double OrbShrinkTscale, OrbShrinkT0;
  double r, phi, M, omega;
  double RK_r, RK_phi, RK_M, RK_omega;
  double Mdot, Macc;
};

struct TimeStep {  // src/Headers/TimeStep.h:10-18 fields
  double t = 0, dt = 0, RK = 0;
  std::vector<int> nri, nzk;
  int Nfr = 0, Nfz = 0;
};

// glibc's additive-feedback rand() seeded with srand(seed), as the reference
// draws it (src/Cell/cell_create_destroy.c:64-71 and the torus initial data,
// src/Cell/initial_data/cell_init_torus.c:31,79): state s[0] = seed,
// s[i] = 16807 s[i-1] mod (2^31 - 1) for i < 31, s[i] = s[i-31] for
// 31 <= i < 34, then s[i] = s[i-31] + s[i-3] mod 2^32; the n-th output is
// s[n+344] >> 1.  The last 34 words are kept in a ring.
struct GlibcRand {
  uint32_t r[34];
  int k = 0;  // r[(k + m) % 34] holds s[n - 34 + m] for the next output n
  explicit GlibcRand(unsigned seed) {
    uint32_t s[344];
    int64_t w = seed == 0 ? 1 : seed;
    s[0] = (uint32_t)w;
    for (int i = 1; i < 31; ++i) {
      w = (16807LL * w) % 2147483647LL;
      if (w < 0) w += 2147483647LL;
      s[i] = (uint32_t)w;
    }
    for (int i = 31; i < 34; ++i) s[i] = s[i - 31];
    for (int i = 34; i < 344; ++i) s[i] = s[i - 31] + s[i - 3];
    for (int m = 0; m < 34; ++m) r[m] = s[310 + m];
  }
  int next() {
    const uint32_t v = r[(k + 3) % 34] + r[(k + 31) % 34];
    r[k] = v;
    k = (k + 1) % 34;
    return (int)(v >> 1);
  }
};

// Per-cell reference handed out by cell_single / face_*_pointer.
struct CellRef;

// Handle behind struct Cell ***.
struct CellsHandle {
  Engine *eng = nullptr;
  Sim *sim = nullptr;
  long long nz = 0;
  int NQ = 5;
  // host mirrors of the device fields (valid when !stale)
  std::vector<double> prim, grad, gradp, tiph, dphi, wiph;
  bool stale = true;       // device is newer than the mirror
  bool dirty_prim = false; // mirror is newer than the device
  bool dirty_tiph = false;
  CellRef *refs = nullptr;
};

struct CellRef {
  CellsHandle *h;
  long long zone;
  // a cell of its own (cell_single_create): NQ primitives, then tiph, dphi,
  // wiph; null for a cell of a Cell *** handle
  double *own = nullptr;
  int NQ = 0;
};

struct Face {  // host copy of one face list (unit-level access)
  CellsHandle *h = nullptr;
  int dir = 0;
  std::vector<int> L, R;
  std::vector<double> r, dphi, dA, deltaL, deltaR, cm;
};

// error reporting shared by all translation units
void disco_set_error(int code, const std::string &msg);

// host geometry (sim_host.cpp) and frame rotation (frame.cpp)
void sim_param_arrays(const Sim *s, DiskParams *dp);
double disco_rOm_a(int type, double r, double a);
double disco_rdrOm_a(int type, double r, double a);
double disco_dtOm_a(int type, double r, double a, double tscale);
// per annulus (k-major, as the annulus table), at separation a: rOm_a,
// rdrOm_a, dtOm_a at the centre and rdrOm_a at the outer face (frame.cpp)
void disco_frame_rows(const Sim *S, double a, std::vector<double> &out);
void disco_cell_offsets(const Sim *s, std::vector<double> &tiph, std::vector<double> &dphi);
