// Types shared by the host engine and the HIP kernels.
//
// Device layout of a batch of D disks (a single simulation is D = 1):
//   * zones of all disks are numbered consecutively; inside a disk,
//     zone(k,i,j) = zone_base + k*NL + aoff[i] + j, NL = sum_i N_p[i]
//     (the reference's cell order, src/Cell/cell_create_destroy.c:41-48);
//   * every per-zone field is its own array (q-major for NQ-vectors);
//   * an "annulus" is one (disk, layer k, ring i); a "pair" is two adjacent
//     annuli that share overlap faces: (k,i)-(k,i+1) in r, (k,i)-(k+1,i) in z.
#pragma once
#include <stdint.h>

#define DISCO_MAXQ 8
#define DISCO_TILE 256

enum { RHO = 0, PPP = 1, URR = 2, UPP = 3, UZZ = 4 };
enum { DDD = 0, TAU = 1, SRR = 2, LLL = 3, SZZ = 4 };
enum { C_FIXED, C_WCELL, C_RIGID, C_KEPLER, C_OMEGA20, C_MILOS, C_POWERLAW };
enum { A_FIXED, A_KEPLER, A_OMEGA20, A_MILOS, A_POWERLAW };
enum { BOUND_FIXED, BOUND_OUTFLOW, BOUND_PERIODIC };
enum { HLLC_SOLVER = 0, HLL_SOLVER = 1 };
// InitialDataType, src/Headers/header.h:8
enum { ID_SHEAR, ID_VORTEX, ID_SSTEST, ID_TORUS, ID_MILOS_MACFADYEN, ID_RAD_DOM, ID_MIDDLE };

// Physics paths a compiled kernel variant keeps.  The batch's launch plan
// picks the smallest variant holding every path its disks take; a cleared
// bit removes that path from the code (and its registers), a set bit keeps
// the disk's runtime choice.  Every variant does the same arithmetic on the
// paths it keeps.
enum {
  F_VISC = 1,   // explicit viscosity (riemann_visc_flux, cell_add_visc_src, viscous CFL)
  F_HLLC = 2,   // RiemannSolver 0
  F_HLL = 4,    // RiemannSolver 1
  F_SINK = 8,   // RhoSinkOn
  F_COOL = 16,  // Cooling
  F_FRAME = 32, // w_analytic_type != 0 (frame rotation terms, split fictitious forces)
  F_ALL = 63
};

// sticky device error codes
enum {
  DERR_NONE = 0,
  DERR_FACE_OVERFLOW = 1,   // more overlap faces than the pair's capacity
  DERR_VISC_VZ = 2,         // viscosity with |v_z| > 1e-12 (the reference exits)
};

struct DiskParams {
  int NR, NZ, NQ, NC;         // annuli / layers incl. ghosts
  int NgMinR, NgMaxR, NgMinZ, NgMaxZ;
  int NglobalZ;
  int ann_base;               // first annulus of this disk in the annulus table
  int rpair_base, zpair_base; // first pair of this disk
  int geo_r, geo_z;           // offsets into the r_faces / z_faces pools
  long long zone_base;
  int NL;                     // zones per layer
  int itile0, itile1;         // this disk's range in the interior (CFL) tile list
  int utile0, utile1;         // this disk's range in the update tile list
  int w_numeric, w_analytic, riemann, visc_const, grav2d, sink, phi_order, cooling;
  int initial_data_type, grav_type, bound_r, no_inner_bc;
  double gamma, cfl, plm, visc, g_eps, rho_floor, cs_floor, cs_cap, vel_cap;
  double orb_tscale, orb_t0, t_max, porho_r1;
};

// Per-disk scalar state living in device memory.
struct GravState {
  double r, phi, M, omega, RK_r, RK_phi, RK_M, RK_omega, Mdot, Macc;
};
struct DiskScalars {
  double t, dt, RK, pad;
  GravState gm[2];
  unsigned long long dtmin_bits;  // CFL minimum of the step (as ordered bits)
  int err, pad2;
  double mdot_acc[2];
  // launch-uniform terms, recomputed whenever the masses move:
  // (xbh, ybh) = (r cos phi, r sin phi) of each mass, epsn = pow(G_EPS, PHI_ORDER)
  double xbh[2], ybh[2], epsn, pad3;
};

// Per annulus: what the fixed r boundaries (cell_boundary_fixed_r,
// src/Cell/cell_boundary.c:208-232) and the boundary damping (cell_bc_damp,
// src/Cell/cell_misc.c:98-132) need of the initial-data plugin's single-cell
// function.  The parts that depend on the annulus alone are tabulated on the
// host at set-up with the C library the reference uses; the parts that
// depend on the cell's angle are evaluated on the device.
enum { INIT_FIXED = 1, INIT_DAMP = 2, INIT_SET = 4 };
struct InitRow {
  double prim[5];   // rho, P, v_r, omega, v_z (INIT_SET: the plugin writes them)
  double a;         // shear: rOm_a(r, 1) / r
  double c0, c1;    // shear: v0 / sqrt(2 pi nu t0) and 4 nu t0
  double rate;      // damping rate (INIT_DAMP)
  int flags;        // INIT_FIXED: a fixed-boundary ghost annulus
  int pad;
};

struct AnnInfo {
  long long zbase;   // first zone of the annulus
  int N;             // N_p
  int disk, k, i;
  int itile;         // its first tile in the interior (CFL) tile list; -1 for ghost annuli and layers
  int utile;         // its first slot in the update tile list (table order; the slot of a tile
                     // does not depend on the order in which a schedule lists the tiles)
  int pad_;
  double rm, rp, zm, zp;  // face positions bounding the annulus
};

struct PairInfo {
  int ann_in, ann_out;   // inner / outer annulus index
  int disk, dir;         // dir 0: r, 1: z
  long long face_base;   // first face slot
  int cap;               // face capacity
  int run_base;          // first run_start slot
  int run_cap;
  int pad;
  double r, deltaL, deltaR, deltaPerp;  // the reference's build_jloop constants
};

struct PairState {   // written by the face builder each substep
  int count;         // faces emitted
  int jpmin;         // outer cell containing the inner annulus' start edge
  int T;             // number of outer-cell advances (runs - 1)
  int seq;           // 1 if the sequential fallback built this pair
  // 1: both annuli have N cells and every inner cell j has exactly two
  // faces, 2j with outer cell (jpmin + j) mod N and 2j + 1 with the next
  // one (no edges tie); the faces' cells, widths and centres then follow
  // from the two annuli's edges alone (pair_face)
  int reg;
};
