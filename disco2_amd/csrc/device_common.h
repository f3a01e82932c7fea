// Device helpers shared by the forward kernels (kernels.hip) and the
// adjoint kernels (adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "physics.h"

#define TILE DISCO_TILE
// minimum waves per SIMD the register allocation of the two largest kernels
// must allow (occupancy against spills)
#ifndef UPDATE_WAVES
#define UPDATE_WAVES 4
#endif
#ifndef FLUX_WAVES
#define FLUX_WAVES 1
#endif
// the update's variant for HLL without viscosity (the 3D disk's)
#ifndef UPDATE_WAVES_HLL
#define UPDATE_WAVES_HLL UPDATE_WAVES
#endif

__device__ __forceinline__ void set_err(int *err, int code) { atomicCAS(err, 0, code); }

// XCD-aware work order.  Workgroups are dispatched round-robin over the 8
// XCDs (workgroup b runs on XCD b mod 8), each with its own L2.  Tiles are
// listed annulus by annulus, so handing XCD x one contiguous band of tiles
// keeps the neighbouring annuli that a tile's faces reach in the same L2.
__device__ __forceinline__ int xcd_order() {
  const int n = gridDim.x, b = blockIdx.x;
  const int q = n >> 3, r = n & 7, x = b & 7;
  return x * q + min(x, r) + (b >> 3);
}

// the workgroup's tile under the mesh's launch schedule
__device__ __forceinline__ int tile_of(const DevMesh &m) { return m.sched_xcd ? xcd_order() : (int)blockIdx.x; }

__device__ __forceinline__ const DiskParams &disk_of(const DevMesh &m, int d) { return m.dp[d]; }

__device__ __forceinline__ int rpair_of(const DiskParams &P, int k, int i) { return P.rpair_base + k * (P.NR - 1) + i; }
__device__ __forceinline__ int zpair_of(const DiskParams &P, int k, int i) { return P.zpair_base + k * P.NR + i; }
// outside the ghost annuli and layers (the cells cell_mindt visits)
__device__ __forceinline__ bool interior_cell(const DiskParams &P, int i, int k) {
  return i >= P.NgMinR && i < P.NR - P.NgMaxR && k >= P.NgMinZ && k < P.NZ - P.NgMaxZ;
}

// An overlap face: its L and R zones, width and centre angle.
struct FaceGeo {
  long long cL, cR;
  double w, cm;
};
// Face f of a regular pair (PairState::reg: inner annulus of N cells from
// zone zin, outer annulus of N cells from zout), from the two annuli's
// edges with the face builder's arithmetic (jloop_cell): face 2j is inner
// cell j's step A with outer cell jo = (jpmin + j) mod N, face 2j + 1 its
// step C with the next outer cell.  m.tiph, m.dphi: the substep's edges.
__device__ __forceinline__ FaceGeo pair_face(const DevMesh &m, int jpmin, int N, long long zin, long long zout, int f) {
  const int j = f >> 1;
  int jo = jpmin + j + (f & 1);
  if (jo >= N) jo -= N;
  FaceGeo g;
  g.cL = zin + j;
  g.cR = zout + jo;
  const double To = m.tiph[zout + jo];
  if (!(f & 1)) {
    const int jm = j == 0 ? N - 1 : j - 1;
    g.w = wrap_p(To - m.tiph[zin + jm]);
    g.cm = To - .5 * g.w;
  } else {
    const double Ti = m.tiph[zin + j];
    g.w = m.dphi[zout + jo] + wrap_pm(Ti - To);
    g.cm = Ti - .5 * g.w;
  }
  return g;
}
// face g of pair PI (state S): computed for a regular pair, else loaded
__device__ __forceinline__ FaceGeo face_geo(const DevMesh &m, const PairInfo &PI, const PairState &S, long long g) {
  if (S.reg) {
    const AnnInfo Ai = m.ann[PI.ann_in];
    return pair_face(m, S.jpmin, Ai.N, Ai.zbase, m.ann[PI.ann_out].zbase, (int)(g - PI.face_base));
  }
  FaceGeo G;
  G.cL = m.fL[g];
  G.cR = m.fR[g];
  G.w = m.fdphi[g];
  G.cm = m.fcm[g];
  return G;
}

// S = (W_R - W_L)/(deltaL + deltaR) of one face, cell_plm.c:36-49 and 86-99
template <int NQ>
__device__ __forceinline__ void face_slope(const DevMesh &m, const PairInfo &P, long long g, double *S) {
  const int cL = m.fL[g], cR = m.fR[g];
  const double cm = m.fcm[g];
  double pL = clean_pi(m.tiph[cL]) - .5 * m.dphi[cL];
  double pR = clean_pi(m.tiph[cR]) - .5 * m.dphi[cR];
  double dpL = wrap_pm(cm - pL);
  double dpR = wrap_pm(pR - cm);
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double WL = m.prim[q * m.Ntot + cL] + dpL * m.gradp[q * m.Ntot + cL];
    double WR = m.prim[q * m.Ntot + cR] - dpR * m.gradp[q * m.Ntot + cR];
    S[q] = (WR - WL) / (P.deltaR + P.deltaL);
  }
}

// Faces of a pair that have inner cell j (local index) as L, and the runs
// that have outer cell jo as R, in face order.
template <typename F>
__device__ __forceinline__ void faces_as_L(const DevMesh &m, int p, const int *fo, long long z, int j, int Nin,
                                           F f) {
  const PairInfo P = m.pairs[p];
  const PairState S = m.pst[p];
  int f1 = (j + 1 < Nin) ? fo[z + 1] : S.count;
  if (f1 > S.count) f1 = S.count;
  int f0 = fo[z];
  if (f0 < 0) f0 = 0;
  for (int n = f0; n < f1; ++n) f(P, P.face_base + n);
}
template <typename F>
__device__ __forceinline__ void faces_as_R(const DevMesh &m, int p, int jo, int Nout, F f) {
  const PairInfo P = m.pairs[p];
  const PairState S = m.pst[p];
  int t = jo - S.jpmin;
  if (t < 0) t += Nout;
  for (; t <= S.T; t += Nout) {
    int f0 = m.run_start[P.run_base + t];
    int f1 = m.run_start[P.run_base + t + 1];
    for (int n = f0; n < f1; ++n) f(P, P.face_base + n);
  }
}

// analytic frame rotation of an annulus (filled by k_scalar whenever the
// masses move; zero for the non-rotating frame)
template <int FEAT = F_ALL>
__device__ __forceinline__ double4 ann_rotation(const DevMesh &m, const DiskParams &P, int a) {
  if (!(FEAT & F_FRAME) || P.w_analytic == A_FIXED) return make_double4(0.0, 0.0, 0.0, 0.0);
  return m.ann_rot[a];
}

// the viscous flux reads cell-centre fields itself; set its pointers
__device__ __forceinline__ void visc_cells(const DevMesh &m, ViscIn &vi, long long cL, long long cR) {
  vi.prim = m.prim;
  vi.grad = m.grad;
  vi.gradp = m.gradp;
  vi.stride = m.Ntot;
  vi.cL = cL;
  vi.cR = cR;
  vi.tL = clean_pi(m.tiph[cL]);
  vi.tR = clean_pi(m.tiph[cR]);
  vi.dL = m.dphi[cL];
  vi.dR = m.dphi[cR];
}

