// Backward kernels of one substep for the hand-written adjoint (2D, NUM_Q 5,
// non-rotating analytic frame, state-independent mesh motion).
//
// The forward's record holds each substep's input and output state, slopes,
// faces and converted conserved state.  In reverse, per substep: the cell
// update (cons2prim, RK blend, prim2cons, sources, the torque of the loss)
// with the outflow remap's backward gathered by the remap's source cells,
// the phi and r faces (Riemann and viscous parts), the r limiter, the face
// slopes and the phi limiter.  Cells gather what their faces hand them (per
// tile in LDS, or per-face slots): no atomics, every sum in a fixed order,
// so the gradients are the same bits run to run.  Launch-uniform scalars
// (dt, alpha, the masses and their positions) are summed per tile and
// reduced per step.
#include <mutex>

#include "adjoint.h"
#include "adjoint_kernels.h"
#include "device_common.h"

__device__ __forceinline__ void load5(const double *a, long long N, long long c, double *v) {
#pragma unroll
  for (int q = 0; q < 5; ++q) v[q] = a[q * N + c];
}
__device__ __forceinline__ void store5(double *a, long long N, long long c, const double *v) {
#pragma unroll
  for (int q = 0; q < 5; ++q) a[q * N + c] = v[q];
}
__device__ __forceinline__ void add5(double *a, long long N, long long c, const double *v) {
#pragma unroll
  for (int q = 0; q < 5; ++q) a[q * N + c] += v[q];
}

// workgroup sum of nine values per thread; thread k < 9 stores sum k at
// dst[k * stride] (every thread calls)
#define SA_FIELDS 9
__device__ void block_sum9(const double *v0, double *dst, long long stride) {
  __shared__ double sh[SA_FIELDS][1024 / 64];
  double v[SA_FIELDS];
#pragma unroll
  for (int k = 0; k < SA_FIELDS; ++k) v[k] = v0[k];
#pragma unroll
  for (int k = 0; k < SA_FIELDS; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < SA_FIELDS; ++k) sh[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x < SA_FIELDS) {
    double t = 0.0;
    for (int i = 0; i < nw; ++i) t += sh[threadIdx.x][i];
    dst[threadIdx.x * stride] = t;
  }
}
// workgroup sum of the scalar adjoints into slot `slot` of the partial sums
// (every thread calls; row k of a.part holds field k of every slot)
__device__ void reduce_store(const ScalarAdj &s, const AdjMesh &a, int slot) {
  const double v[SA_FIELDS] = {s.dt, s.alpha, s.M[0], s.M[1], s.xbh[0], s.xbh[1], s.ybh[0], s.ybh[1], s.loss};
  block_sum9(v, a.part + slot, a.npart);
}

__device__ __forceinline__ void load_pair(const DevMesh &m, long long cL, long long cR, FacePair &c) {
  load5(m.prim, m.Ntot, cL, c.pL);
  load5(m.prim, m.Ntot, cR, c.pR);
  load5(m.grad, m.Ntot, cL, c.gL);
  load5(m.grad, m.Ntot, cR, c.gR);
  load5(m.gradp, m.Ntot, cL, c.gpL);
  load5(m.gradp, m.Ntot, cR, c.gpR);
}
// A face's width and centre follow the rotations of the annuli whose edges
// bound it (face_create_destroy.c:74-125): d(width)/d(theta_in, theta_out)
// and d(centre)/d(theta_in, theta_out) by face type.  The adjoints of a face's
// width (wb), centre (cb) and of its cells' own edges (tLb, tRb) go to the
// rotation adjoint of the inner (L cell, vin) and outer (R cell, vout) annulus.
__device__ __forceinline__ void face_rot_split(signed char type, double wb, double cb, double tLb, double tRb,
                                               double *vin, double *vout) {
  double w_in = 0.0, w_out = 0.0, c_in = 0.0, c_out = 0.0;
  switch (type) {
    case FACE_II: c_in = 1.0; break;
    case FACE_IO: w_in = -1.0; w_out = 1.0; c_in = c_out = 0.5; break;
    case FACE_OO: c_out = 1.0; break;
    default: w_in = 1.0; w_out = -1.0; c_in = c_out = 0.5; break;  // FACE_OI
  }
  *vin = wb * w_in + cb * c_in + tLb;
  *vout = wb * w_out + cb * c_out + tRb;
}

// Outflow remap backward (k_bc_r, cell_boundary_outflow_r), gathered.  The
// remap fills ghost annulus NR-2 from the faces' overlap with NR-3, then
// NR-1 from the new NR-2 (and annulus 0 from annulus 1 unless NoInnerBC):
// ghost = sum over its faces of prim[src] * dA_f / dAc, dA_f = r * width *
// dperp, dAc the ghost's face area.  Backward, each source cell gathers
// what its ghosts hand it; a cell of NR-3 first adds to each NR-2 ghost the
// NR-1 ghosts' shares.  A face's width also enters its dA: the adjoint of
// the width goes to the rotations of the face's two annuli, and each of its
// two cells takes its own side's share (face_rot_split), so no cell is
// written by another's thread.  The mesh view has the substep's faces and
// the new dphi.
// dAc of ghost cell g of annulus An (inner: the inner ghost annulus)
__device__ __forceinline__ double ghost_area(const DevMesh &m, const AnnInfo &An, long long g, bool inner) {
  return (An.zp - An.zm) * (inner ? An.rp : An.rm) * m.dphi_new[g];
}
// the adjoint ghost cell g of annulus NR-2 (An, local index jg) holds after
// the NR-1 ghosts' shares (the second stage's backward)
__device__ __forceinline__ void ghost2_total(const DevMesh &m, const AdjMesh &a, const DiskParams &P, int k,
                                             const AnnInfo &An, const AnnInfo &Ao, long long g, int jg, double *lam) {
  load5(a.aPn, m.Ntot, g, lam);
  faces_as_L(m, rpair_of(P, k, P.NR - 2), m.fo_r, g, jg, An.N, [&](const PairInfo &PI, long long f) {
    const long long h = m.fR[f];
    const double w = PI.r * m.fdphi[f] * PI.deltaPerp / ghost_area(m, Ao, h, false);
    for (int q = 0; q < 5; ++q) lam[q] += a.aPn[q * m.Ntot + h] * w;
  });
}
// the rotation adjoint of the L (asL) or R side of face f whose width
// adjoint is wb
__device__ __forceinline__ double side_rot(const DevMesh &m, long long f, double wb, bool asL) {
  double vin, vout;
  face_rot_split(m.ftype[f], wb, 0.0, 0.0, 0.0, &vin, &vout);
  return asL ? vin : vout;
}
// cell c (local j) of annulus A: the remap's backward gathered into ob (a
// source cell) and its sides' rotation adjoints of the remap's faces into
// *rot (source and ghost cells)
__device__ __forceinline__ void remap_gather(const DevMesh &m, const AdjMesh &a, const DiskParams &P,
                                             const int2 tl, const AnnInfo &A, long long c, int j, double *ob,
                                             double *rot) {
  const int k = A.k, i = A.i, NR = P.NR;
  const bool inner = P.no_inner_bc != 1;
  double rt = 0.0;
  if (i == NR - 3) {  // source of the NR-2 ghosts (with the NR-1 ghosts' shares)
    const AnnInfo An = m.ann[tl.x + 1], Ao = m.ann[tl.x + 2];
    faces_as_L(m, rpair_of(P, k, NR - 3), m.fo_r, c, j, A.N, [&](const PairInfo &PI, long long f) {
      const long long g = m.fR[f];
      double lam[5];
      ghost2_total(m, a, P, k, An, Ao, g, (int)(g - An.zbase), lam);
      const double s = PI.r * PI.deltaPerp / ghost_area(m, An, g, false);
      const double w = m.fdphi[f] * s;
      double wb = 0.0;
      for (int q = 0; q < 5; ++q) {
        ob[q] += lam[q] * w;
        wb += lam[q] * m.prim_new[q * m.Ntot + c];
      }
      rt += side_rot(m, f, wb * s, true);
    });
  }
  if (i == NR - 2) {
    // source of the NR-1 ghosts
    const AnnInfo Ao = m.ann[tl.x + 1];
    faces_as_L(m, rpair_of(P, k, NR - 2), m.fo_r, c, j, A.N, [&](const PairInfo &PI, long long f) {
      const long long h = m.fR[f];
      const double s = PI.r * PI.deltaPerp / ghost_area(m, Ao, h, false);
      double wb = 0.0;
      for (int q = 0; q < 5; ++q) wb += a.aPn[q * m.Ntot + h] * m.prim_new[q * m.Ntot + c];
      rt += side_rot(m, f, wb * s, true);
    });
    // ghost of the NR-3 cells
    double lam[5];
    ghost2_total(m, a, P, k, A, Ao, c, j, lam);
    const double s0 = (A.zp - A.zm) * A.rm * m.dphi_new[c];
    faces_as_R(m, rpair_of(P, k, NR - 3), j, A.N, [&](const PairInfo &PI, long long f) {
      const long long src = m.fL[f];
      double wb = 0.0;
      for (int q = 0; q < 5; ++q) wb += lam[q] * m.prim_new[q * m.Ntot + src];
      rt += side_rot(m, f, wb * (PI.r * PI.deltaPerp / s0), false);
    });
  }
  if (i == NR - 1) {  // ghost of the NR-2 cells
    double lam[5];
    load5(a.aPn, m.Ntot, c, lam);
    const double s0 = (A.zp - A.zm) * A.rm * m.dphi_new[c];
    faces_as_R(m, rpair_of(P, k, NR - 2), j, A.N, [&](const PairInfo &PI, long long f) {
      const long long src = m.fL[f];
      double wb = 0.0;
      for (int q = 0; q < 5; ++q) wb += lam[q] * m.prim_new[q * m.Ntot + src];
      rt += side_rot(m, f, wb * (PI.r * PI.deltaPerp / s0), false);
    });
  }
  if (i == 1 && inner) {  // source of the inner ghosts (NR = 4: as well as the outer)
    const AnnInfo Ai = m.ann[tl.x - 1];
    faces_as_R(m, rpair_of(P, k, 0), j, A.N, [&](const PairInfo &PI, long long f) {
      const long long g = m.fL[f];
      const double s = PI.r * PI.deltaPerp / ghost_area(m, Ai, g, true);
      const double w = m.fdphi[f] * s;
      double wb = 0.0;
      for (int q = 0; q < 5; ++q) {
        const double lam = a.aPn[q * m.Ntot + g];
        ob[q] += lam * w;
        wb += lam * m.prim_new[q * m.Ntot + c];
      }
      rt += side_rot(m, f, wb * s, false);
    });
  }
  if (i == 0 && inner) {  // inner ghost of the annulus-1 cells
    double lam[5];
    load5(a.aPn, m.Ntot, c, lam);
    const double s0 = (A.zp - A.zm) * A.rp * m.dphi_new[c];
    faces_as_L(m, rpair_of(P, k, 0), m.fo_r, c, j, A.N, [&](const PairInfo &PI, long long f) {
      const long long src = m.fR[f];
      double wb = 0.0;
      for (int q = 0; q < 5; ++q) wb += lam[q] * m.prim_new[q * m.Ntot + src];
      rt += side_rot(m, f, wb * (PI.r * PI.deltaPerp / s0), true);
    });
  }
  *rot = rt;
}

// Cell update backward (k_update phase 2): cons2prim, the RK blend and the
// RK snapshot, prim2cons, the sources and the torque term.  Writes the
// adjoint of the cell's conserved state (au) and starts aP.
// minimum waves per SIMD of the update and viscous-part kernels (1: no bound)
#ifndef ADJ_UPDATE_WAVES
#define ADJ_UPDATE_WAVES 1
#endif
#ifndef ADJ_VISC_WAVES
#define ADJ_VISC_WAVES 1
#endif
__device__ void rot_annulus(const DevMesh &m, const double *src, int ann, double *out);
__device__ void pick_cell(const DevMesh &m, const int2 *itiles, int nitiles, int nann,
                          const unsigned long long *dtp, CflGrad *pick);
// Workgroups past the nmain tiles: one per annulus of the rotation sums,
// then (x.pick) one for the step's CFL cell.
__global__ __launch_bounds__(TILE, ADJ_UPDATE_WAVES) void ka_update(DevMesh m, AdjMesh a, const int2 *tiles,
                                                                    int nmain, int nrot, const int2 *itiles,
                                                                    int nitiles, int nann, AdjExtra x) {
  const int ti = xcd_order();
  if (ti >= nmain + nrot) {
    DevMesh mp = m;  // the step-start state and the scalars its CFL minimum saw
    mp.prim = (double *)x.pick_prim;
    mp.tiph = x.pick_tiph;
    mp.dphi = (double *)x.pick_dphi;
    mp.sc = (DiskScalars *)x.pick_sc;
    mp.dt_tile = (double *)x.pick_dt_tile;
    pick_cell(mp, itiles, nitiles, nann, x.dtbits, x.pick);
    return;
  }
  if (ti >= nmain) {
    rot_annulus(m, x.rot_src, ti - nmain, x.rot_out);
    return;
  }
  const int2 tl = tiles[ti];
  const AnnInfo A = m.ann[tl.x];
  const DiskParams &P = disk_of(m, A.disk);
  const DiskScalars &sc = m.sc[A.disk];
  const int j = tl.y + threadIdx.x;
  ScalarAdj sb = {};
  if (j < A.N) {
    const long long c = A.zbase + j;
    const double rp = A.rp, rm = A.rm, zp = A.zp, zm = A.zm;
    const double r = .5 * (rp + rm), dz = zp - zm;
    const double dphi = m.dphi[c], dphin = m.dphi_new[c];
    double pr[5], u[5], ob[5] = {0, 0, 0, 0, 0}, ub[5] = {0, 0, 0, 0, 0}, prb[5] = {0, 0, 0, 0, 0}, u0b[5];
    load5(m.prim, m.Ntot, c, pr);
    load5(m.u_save, m.Ntot, c, u);
    // the remap overwrites the ghost annuli: their adjoint goes to the
    // cells they were averaged from, which gather it
    const int i = A.i, NR = P.NR;
    const bool ghost = i >= NR - 2 || (i == 0 && P.no_inner_bc != 1);
    if (!ghost) load5(a.aPn, m.Ntot, c, ob);
    double rot = 0.0;
    if (i >= NR - 3 || i <= 1) remap_gather(m, a, P, tl, A, c, j, ob, &rot);
    vjp_cons2prim(u, r, .5 * (rp * rp - rm * rm) * dphin * dz, P, ob, ub);
    store5(a.au, m.Ntot, c, ub);
    for (int q = 0; q < 5; ++q) {
      if (a.copy_first) {
        u0b[q] = ub[q] + a.aR[q * m.Ntot + c];  // u = u0 (RK = 0) and RKcons = u0
      } else {
        u0b[q] = (1.0 - sc.RK) * ub[q];
        a.aR[q * m.Ntot + c] = sc.RK * ub[q];
      }
    }
    vjp_prim2cons(pr, r, .5 * (rp * rp - rm * rm) * dphi * dz, P.gamma, u0b, prb);
    const double phi = clean_pi(m.tiph[c]) - .5 * dphi;
    const double dVs = dphi * .5 * (rp * rp - rm * rm) * dz;
    const double wl = interior_cell(P, A.i, A.k) ? a.wl : 0.0;
    double phib = 0.0;
    vjp_cell_sources(P, sc, pr, r, dVs, a.fac * sc.dt, phi, .5 * (zp + zm), ub, wl, prb, sb, &phib);
    a.arot[c] += rot + phib;
    store5(a.aP, m.Ntot, c, prb);
  }
  reduce_store(sb, a, ti);
}

// phi face (j, j+1): cell j+1 received +X (+Y), cell j -X.
struct PhiFaceAt {
  long long cL, cR;
  double Xb[5], Yb[3];
  PhiFaceGeo g;
};
__device__ __forceinline__ void phi_face_at(const DevMesh &m, const AdjMesh &a, const AnnInfo &A, int ann,
                                            const DiskParams &P, int j, PhiFaceAt &F) {
  F.cL = A.zbase + j;
  F.cR = A.zbase + (j == A.N - 1 ? 0 : j + 1);
  const double r = .5 * (A.rp + A.rm);
  for (int q = 0; q < 5; ++q) F.Xb[q] = a.au[q * m.Ntot + F.cR] - a.au[q * m.Ntot + F.cL];
  F.Yb[0] = a.au[SRR * m.Ntot + F.cR] / r - a.au[SRR * m.Ntot + F.cL] / r;
  F.Yb[1] = a.au[LLL * m.Ntot + F.cR] - a.au[LLL * m.Ntot + F.cL];
  F.Yb[2] = a.au[TAU * m.Ntot + F.cR] - a.au[TAU * m.Ntot + F.cL];
  F.g.r = r;
  F.g.dA = (A.rp - A.rm) * (A.zp - A.zm);
  F.g.w = m.wiph[ann];
  F.g.tL = clean_pi(m.tiph[F.cL]);
  F.g.tR = clean_pi(m.tiph[F.cR]);
  F.g.dL = m.dphi[F.cL];
  F.g.dR = m.dphi[F.cR];
  F.g.rdr = ann_rotation(m, P, ann).y;
}

// The phi-face backward kernels run over tiles of TILE - 1 cells j0 .. j0+n-1:
// thread t solves face slot t = (j0-1+t, j0+t) for t = 0..n, keeps what each
// of its two cells receives in LDS, and thread t then adds slot t (its cell
// is R) and slot t+1 (its cell is L) into its own cell: no atomics.  A face
// on a tile boundary is solved by both tiles; only slots 0..n-1 add to the
// scalar adjoints, which counts every face once.

// phi faces backward, Riemann part
// minimum waves per SIMD of the Riemann-part face kernels (1: no bound)
#ifndef ADJ_HYDRO_WAVES
#define ADJ_HYDRO_WAVES 1
#endif
__device__ __forceinline__ void phi_h_tile(const DevMesh &m, const AdjMesh &a, const int2 *tiles, int ti, int slot0) {
  const int2 tl = tiles[ti];
  const AnnInfo A = m.ann[tl.x];
  const DiskParams &P = disk_of(m, A.disk);
  const int t = threadIdx.x;
  const int n = min(TILE - 1, A.N - tl.y);
  __shared__ double sL[5][TILE], sR[5][TILE];
  ScalarAdj sb = {};
  if (t <= n) {
    int jf = tl.y - 1 + t;
    if (jf < 0) jf += A.N;
    PhiFaceAt F;
    phi_face_at(m, a, A, tl.x, P, jf, F);
    FacePair c;
    load_pair(m, F.cL, F.cR, c);
    double pLb[5], pRb[5];
    ScalarAdj fs = {};
    vjp_phiface_hydro(P, c, F.g, a.fac * m.sc[A.disk].dt, F.Xb, fs, pLb, pRb);
    if (t < n) sb = fs;
    for (int q = 0; q < 5; ++q) {
      sL[q][t] = pLb[q];
      sR[q][t] = pRb[q];
    }
  }
  __syncthreads();
  if (t < n) {
    const long long c = A.zbase + tl.y + t;
    const double dph = m.dphi[c];
    for (int q = 0; q < 5; ++q) {
      const double bR = sR[q][t], bL = sL[q][t + 1];
      a.aP[q * m.Ntot + c] += bR + bL;
      a.agradp[q * m.Ntot + c] = .5 * dph * bL - .5 * dph * bR;  // first writer of agradp
    }
  }
  reduce_store(sb, a, slot0 + ti);
}

// phi faces backward, viscous part
__device__ __forceinline__ void phi_v_tile(const DevMesh &m, const AdjMesh &a, const int2 *tiles, int ti, int slot0) {
  const int2 tl = tiles[ti];
  const AnnInfo A = m.ann[tl.x];
  const DiskParams &P = disk_of(m, A.disk);
  const DiskScalars &sc = m.sc[A.disk];
  const int t = threadIdx.x;
  const int n = min(TILE - 1, A.N - tl.y);
  // per slot: Avgb/2 (5), Gphb/(r dphi) (2), Grrb/2 (2), the face angle's adjoint
  __shared__ double sv[10][TILE];
  ScalarAdj sb = {};
  if (t <= n) {
    int jf = tl.y - 1 + t;
    if (jf < 0) jf += A.N;
    PhiFaceAt F;
    phi_face_at(m, a, A, tl.x, P, jf, F);
    FacePair c;
    load_pair(m, F.cL, F.cR, c);
    PhiFaceViscAdj v;
    ScalarAdj fs = {};
    vjp_phiface_visc(P, sc, c, F.g, a.fac * sc.dt, F.Yb, fs, v);
    if (t < n) sb = fs;
    const double dt_phi = (F.g.tR - F.g.tL) * F.g.r;
    for (int q = 0; q < 5; ++q) sv[q][t] = .5 * v.Avgb[q];
    for (int k = 0; k < 2; ++k) {
      sv[5 + k][t] = v.Gphb[k] / dt_phi;
      sv[7 + k][t] = .5 * v.Grrb[k];
    }
    sv[9][t] = v.tLb;
  }
  __syncthreads();
  if (t < n) {
    const long long c = A.zbase + tl.y + t;
    for (int q = 0; q < 5; ++q) a.aP[q * m.Ntot + c] += sv[q][t] + sv[q][t + 1];
    for (int k = 0; k < 2; ++k) {
      a.aP[(URR + k) * m.Ntot + c] += sv[5 + k][t] - sv[5 + k][t + 1];
      a.agrad[(URR + k) * m.Ntot + c] = sv[7 + k][t] + sv[7 + k][t + 1];  // read by ka_plm_r
    }
    a.arot[c] += sv[9][t + 1];
  }
  reduce_store(sb, a, slot0 + ti);
}

// r face g of a pair: the R cell received +X (+Y/r_R), the L cell -X.
struct RFaceAt {
  long long g, cL, cR;
  double Xb[5], Yb[3];
  RFaceGeo geo;
  double wscale;  // d(dA)/d(face width): r deltaPerp
};
__device__ __forceinline__ void r_face_at(const DevMesh &m, const AdjMesh &a, const PairInfo &PI,
                                          const DiskParams &P, long long g, RFaceAt &F) {
  F.g = g;
  F.cL = m.fL[g];
  F.cR = m.fR[g];
  for (int q = 0; q < 5; ++q) F.Xb[q] = a.au[q * m.Ntot + F.cR] - a.au[q * m.Ntot + F.cL];
  F.Yb[0] = a.au[SRR * m.Ntot + F.cR] / (PI.r + PI.deltaR) - a.au[SRR * m.Ntot + F.cL] / (PI.r - PI.deltaL);
  F.Yb[1] = a.au[LLL * m.Ntot + F.cR] - a.au[LLL * m.Ntot + F.cL];
  F.Yb[2] = a.au[TAU * m.Ntot + F.cR] - a.au[TAU * m.Ntot + F.cL];
  F.geo.r = PI.r;
  F.geo.deltaL = PI.deltaL;
  F.geo.deltaR = PI.deltaR;
  F.geo.dA = PI.r * m.fdphi[g] * PI.deltaPerp;
  F.geo.cm = m.fcm[g];
  F.geo.tL = clean_pi(m.tiph[F.cL]);
  F.geo.dL = m.dphi[F.cL];
  F.geo.tR = clean_pi(m.tiph[F.cR]);
  F.geo.dR = m.dphi[F.cR];
  F.geo.rdr = ann_rotation(m, P, PI.ann_in).w;
  F.wscale = PI.r * PI.deltaPerp;
}

// The r-face backward kernels write what each face hands its two cells into
// per-face slots of a.fr (q-major over faces, AFR_* offsets in units of
// NFtot); ka_plm_r gathers them per cell over the cell's face ranges.
__device__ __forceinline__ void rflux_h_tile(const DevMesh &m, const AdjMesh &a, const int2 *ftiles, int ti, int slot0) {
  const int2 tl = ftiles[ti];
  const int p = tl.x;
  const int f = tl.y + threadIdx.x;
  ScalarAdj sb = {};
  if (f < m.pst[p].count) {
    const PairInfo PI = m.pairs[p];
    const DiskParams &P = disk_of(m, PI.disk);
    RFaceAt F;
    r_face_at(m, a, PI, P, PI.face_base + f, F);
    FacePair c;
    load_pair(m, F.cL, F.cR, c);
    RFaceHydroAdj h;
    vjp_rface_hydro(P, c, F.geo, a.fac * m.sc[PI.disk].dt, F.Xb, sb, h);
    const long long NF = m.NFtot, g = F.g;
    {  // the face slope S, stored for ka_plm_r's replay of the limiter
      double S[5];
      face_slope<5>(m, PI, g, S);
      for (int q = 0; q < 5; ++q) m.Fface[q * NF + g] = S[q];
    }
    for (int q = 0; q < 5; ++q) {
      a.fr[(AFR_HL + q) * NF + g] = h.pLb[q];
      a.fr[(AFR_HR + q) * NF + g] = h.pRb[q];
    }
    // dpL = cm - (tL - dL/2), dpR = (tR - dR/2) - cm
    double vin, vout;
    face_rot_split(m.ftype[g], h.dAb * F.wscale, h.dpLb - h.dpRb, -h.dpLb, h.dpRb, &vin, &vout);
    a.fr[AFR_ROT * NF + g] = vin;
    a.fr[(AFR_ROT + 1) * NF + g] = vout;
  }
  reduce_store(sb, a, slot0 + ti);
}

__device__ __forceinline__ void rflux_v_tile(const DevMesh &m, const AdjMesh &a, const int2 *ftiles, int ti, int slot0) {
  const int2 tl = ftiles[ti];
  const int p = tl.x;
  const int f = tl.y + threadIdx.x;
  ScalarAdj sb = {};
  if (f < m.pst[p].count) {
    const PairInfo PI = m.pairs[p];
    const DiskParams &P = disk_of(m, PI.disk);
    const DiskScalars &sc = m.sc[PI.disk];
    RFaceAt F;
    r_face_at(m, a, PI, P, PI.face_base + f, F);
    FacePair c;
    load_pair(m, F.cL, F.cR, c);
    RFaceViscAdj v;
    vjp_rface_visc(P, sc, c, F.geo, a.fac * sc.dt, F.Yb, sb, v);
    const long long NF = m.NFtot, g = F.g;
    for (int q = 0; q < 5; ++q) {
      a.fr[(AFR_WL + q) * NF + g] = v.WLb[q];
      a.fr[(AFR_WR + q) * NF + g] = v.WRb[q];
    }
    for (int k = 0; k < 2; ++k) a.fr[(AFR_G + k) * NF + g] = .5 * v.Gphb[k] / F.geo.r;
    double vin, vout;
    face_rot_split(m.ftype[g], v.dAb * F.wscale, v.cmb + v.dpLb - v.dpRb, -v.dpLb, v.dpRb, &vin, &vout);
    a.fr[AFR_ROT * NF + g] += vin;
    a.fr[(AFR_ROT + 1) * NF + g] += vout;
  }
  reduce_store(sb, a, slot0 + ti);
}

// The phi faces' and the r faces' backward of one part (Riemann or viscous)
// in one launch: workgroups [0, nu) take the phi-face tiles, the rest the
// r-face tiles.  They are independent (the phi faces add into the cells'
// adjoints, the r faces write per-face slots) and their scalar partials go
// to consecutive slots from slot0, as two launches would write them.
__global__ __launch_bounds__(TILE, ADJ_HYDRO_WAVES) void ka_faces_h(DevMesh m, AdjMesh a, const int2 *utiles, int nu,
                                                                    const int2 *ftiles, int slot0) {
  const int b = xcd_order();
  if (b < nu)
    phi_h_tile(m, a, utiles, b, slot0);
  else
    rflux_h_tile(m, a, ftiles, b - nu, slot0 + nu);
}
__global__ __launch_bounds__(TILE, ADJ_VISC_WAVES) void ka_faces_v(DevMesh m, AdjMesh a, const int2 *utiles, int nu, const int2 *ftiles,
                                                   int slot0) {
  const int b = xcd_order();
  if (b < nu)
    phi_v_tile(m, a, utiles, b, slot0);
  else
    rflux_v_tile(m, a, ftiles, b - nu, slot0 + nu);
}

// r slope backward (k_plm_rz, r direction).  The cell gathers what its r
// faces' flux adjoints hand it (primitives, r and phi slopes, its annulus
// rotation) in the same walk that replays the slope sum; then it replays
// the sequential limiter (on the face slopes rflux_h_tile left in Fface) and
// hands each face its share of the slope's adjoint: dA/dAtot to every face
// when the sum survived, PLM to the face whose S the limiter took last,
// nothing when it was zeroed.  Each face gets one value from its L cell
// (aSL, and the adjoint of its width AFR_WB) and one from its R cell (aSR,
// AFR_WB + 1).
__global__ __launch_bounds__(TILE) void ka_plm_r(DevMesh m, AdjMesh a, const int2 *tiles) {
  const int2 tl = tiles[xcd_order()];
  const AnnInfo A = m.ann[tl.x];
  const int j = tl.y + threadIdx.x;
  if (j >= A.N) return;
  const DiskParams &P = disk_of(m, A.disk);
  const long long z = A.zbase + j;
  const long long NF = m.NFtot;
  const int k = A.k, i = A.i;
  const int pR = i >= 1 ? rpair_of(P, k, i - 1) : -1;
  const int pLf = i <= P.NR - 2 ? rpair_of(P, k, i) : -1;
  const bool visc = a.visc;
  const double tc = clean_pi(m.tiph[z]), dc = m.dphi[z];
  // the cell's faces in the walk order of k_plm_rz: as R, its runs in the
  // pair below (at most two, the second is the wrap), then as L one range
  // in the pair above; range u is faces [fb[u] + b[u], fb[u] + e[u])
  long long fb[3] = {0, 0, 0};
  int b[3] = {0, 0, 0}, e[3] = {0, 0, 0};
  double pr[3] = {0, 0, 0}, pp[3] = {0, 0, 0}, del[3] = {0, 0, 0};
  if (pR >= 0) {
    const PairInfo PI = m.pairs[pR];
    const PairState S = m.pst[pR];
    int t = j - S.jpmin;
    if (t < 0) t += A.N;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (t <= S.T) {
        fb[u] = PI.face_base;
        b[u] = m.run_start[PI.run_base + t];
        e[u] = m.run_start[PI.run_base + t + 1];
        pr[u] = PI.r;
        pp[u] = PI.deltaPerp;
        del[u] = PI.deltaR;
      }
      t += A.N;
    }
  }
  if (pLf >= 0) {
    const PairInfo PI = m.pairs[pLf];
    const PairState S = m.pst[pLf];
    int f1 = (j + 1 < A.N) ? m.fo_r[z + 1] : S.count;
    if (f1 > S.count) f1 = S.count;
    int f0 = m.fo_r[z];
    if (f0 < 0) f0 = 0;
    fb[2] = PI.face_base;
    b[2] = f0;
    e[2] = f1;
    pr[2] = PI.r;
    pp[2] = PI.deltaPerp;
    del[2] = PI.deltaL;
  }
  double ap[5] = {0, 0, 0, 0, 0}, ag[5] = {0, 0, 0, 0, 0}, agp[5] = {0, 0, 0, 0, 0}, rot = 0.0;
  double g[5] = {0, 0, 0, 0, 0};
  // the face slopes of the cell's first NC faces wait in LDS (a private
  // column per thread, no barrier) for the limiter replay and the shares
  constexpr int NC = 4;
  __shared__ double s_S[NC][5][TILE];
  const int t = threadIdx.x;
  int n = 0;
  // the slope sum and the gather.  As R: primR = p - g deltaR - gp dpR
  // (Riemann part), W_R = p - dpR gp (viscous); as L: primL = p + g deltaL
  // + gp dpL, W_L = p + dpL gp
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    const bool asL = u == 2;
    for (int f = b[u]; f < e[u]; ++f, ++n) {
      const long long gi = fb[u] + f;
      const double dA = pr[u] * m.fdphi[gi] * pp[u];
      double S[5];
#pragma unroll
      for (int q = 0; q < 5; ++q) S[q] = m.Fface[q * NF + gi];
#pragma unroll
      for (int q = 0; q < 5; ++q) g[q] += S[q] * dA;
      if (n < NC)
#pragma unroll
        for (int q = 0; q < 5; ++q) s_S[n][q][t] = S[q];
      const double dpo = asL ? wrap_pm(m.fcm[gi] - (tc - .5 * dc)) : wrap_pm(-m.fcm[gi] + (tc - .5 * dc));
      const int hs = asL ? AFR_HL : AFR_HR, ws = asL ? AFR_WL : AFR_WR;
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        const double h = a.fr[(hs + q) * NF + gi];
        const double sv = visc ? h + a.fr[(ws + q) * NF + gi] : h;
        ap[q] += sv;
        if (asL) {
          ag[q] += del[u] * h;
          agp[q] += dpo * sv;
        } else {
          ag[q] -= del[u] * h;
          agp[q] -= dpo * sv;
        }
      }
      if (visc)
        for (int kk = 0; kk < 2; ++kk) agp[URR + kk] += a.fr[(AFR_G + kk) * NF + gi];
      rot += a.fr[(asL ? AFR_ROT : AFR_ROT + 1) * NF + gi];
    }
  }
  // the r slope's adjoint: its r faces' plus (viscous) its phi faces' (agrad)
  double gb[5];
  for (int q = 0; q < 5; ++q) {
    a.aP[q * m.Ntot + z] += ap[q];
    gb[q] = ag[q];
    a.agradp[q * m.Ntot + z] += agp[q];
  }
  if (visc)
    for (int kk = 0; kk < 2; ++kk) gb[URR + kk] += a.agrad[(URR + kk) * m.Ntot + z];
  a.arot[z] += rot;
  const double dz = A.zp - A.zm, dp = m.dphi[z];
  double dAtot;
  if (i == 0)
    dAtot = dz * A.rp * dp;
  else if (i == P.NR - 1)
    dAtot = dz * A.rm * dp;
  else
    dAtot = dz * (A.rp + A.rm) * dp;
  for (int q = 0; q < 5; ++q) g[q] /= dAtot;
  // replay of the sequential limiter: which face's S it took last
  const double PLM = P.plm;
  int src[5] = {-1, -1, -1, -1, -1};
  auto slope = [&](int nn, long long gi, double *S) {
    if (nn < NC) {
#pragma unroll
      for (int q = 0; q < 5; ++q) S[q] = s_S[nn][q][t];
    } else {
#pragma unroll
      for (int q = 0; q < 5; ++q) S[q] = m.Fface[q * NF + gi];
    }
  };
  n = 0;
#pragma unroll
  for (int u = 0; u < 3; ++u)
    for (int f = b[u]; f < e[u]; ++f) {
      double S[5];
      slope(n, fb[u] + f, S);
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        if (S[q] * g[q] < 0.0) {
          g[q] = 0.0;
          src[q] = -2;
        } else if (fabs(PLM * S[q]) < fabs(g[q])) {
          g[q] = PLM * S[q];
          src[q] = n;
        }
      }
      ++n;
    }
  // each face's share of the slope's adjoint: dA/dAtot to every face when the
  // sum survived, PLM to the face the limiter took last, nothing when zeroed;
  // and the adjoint of the face's width (AFR_WB from the L cell, + 1 from R)
  n = 0;
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    double *dst = u == 2 ? a.aSL : a.aSR;
    for (int f = b[u]; f < e[u]; ++f) {
      const long long gi = fb[u] + f;
      const double dA = pr[u] * m.fdphi[gi] * pp[u];
      double S[5];
      slope(n, gi, S);
      double wb = 0.0;
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        dst[q * NF + gi] = src[q] == -1 ? dA / dAtot * gb[q] : (src[q] == n ? PLM * gb[q] : 0.0);
        if (src[q] == -1) wb += gb[q] * S[q] / dAtot;  // d g / d dA_f
      }
      a.fr[(AFR_WB + (u == 2 ? 0 : 1)) * NF + gi] = wb * pr[u] * pp[u];
      ++n;
    }
  }
}

// face slope S = (W_R - W_L)/(deltaL + deltaR), W_L = p_L + dpL gp_L,
// W_R = p_R - dpR gp_R, backward, gathered per cell over its r faces: cell j
// (zone z) of annulus A receives ap (its primitives), agp (its phi slopes)
// and rot (its annulus rotation)
__device__ __forceinline__ void slope_cell(const DevMesh &m, const AdjMesh &a, const AnnInfo &A, const DiskParams &P,
                                           int j, long long z, double *ap, double *agp, double *rot_out) {
  const long long NF = m.NFtot;
  const int k = A.k, i = A.i;
  const int pR = i >= 1 ? rpair_of(P, k, i - 1) : -1;
  const int pLf = i <= P.NR - 2 ? rpair_of(P, k, i) : -1;
  const double tc = clean_pi(m.tiph[z]), dc = m.dphi[z];
  double gpc[5], rot = 0.0;
  for (int q = 0; q < 5; ++q) ap[q] = agp[q] = 0.0;
  load5(m.gradp, m.Ntot, z, gpc);
  // one face: Sb the slope's adjoint, W_Lb = -Sb/d, W_Rb = Sb/d; dpLb, dpRb
  // the adjoints of the two offsets, which move the face centre and edges
  auto face = [&](const PairInfo &PI, long long gi, bool asL) {
    const long long o = asL ? m.fR[gi] : m.fL[gi];
    const double cm = m.fcm[gi];
    const double d = PI.deltaR + PI.deltaL;
    const double dpc = asL ? wrap_pm(cm - (tc - .5 * dc)) : wrap_pm((tc - .5 * dc) - cm);
    double dpLb = 0.0, dpRb = 0.0;
    for (int q = 0; q < 5; ++q) {
      const double Sb = a.aSL[q * NF + gi] + a.aSR[q * NF + gi];
      const double gpo = m.gradp[q * m.Ntot + o];
      if (asL) {
        const double WLb = -Sb / d;
        ap[q] += WLb;
        agp[q] += dpc * WLb;
        dpLb += WLb * gpc[q];
        dpRb -= Sb / d * gpo;
      } else {
        const double WRb = Sb / d;
        ap[q] += WRb;
        agp[q] -= dpc * WRb;
        dpLb += -Sb / d * gpo;
        dpRb -= WRb * gpc[q];
      }
    }
    const double wb = a.fr[AFR_WB * NF + gi] + a.fr[(AFR_WB + 1) * NF + gi];
    double vin, vout;
    // dpL = cm - (tL - dL/2), dpR = (tR - dR/2) - cm
    face_rot_split(m.ftype[gi], wb, dpLb - dpRb, -dpLb, dpRb, &vin, &vout);
    rot += asL ? vin : vout;
  };
  if (pR >= 0) faces_as_R(m, pR, j, A.N, [&](const PairInfo &PI, long long gi) { face(PI, gi, false); });
  if (pLf >= 0) faces_as_L(m, pLf, m.fo_r, z, j, A.N, [&](const PairInfo &PI, long long gi) { face(PI, gi, true); });
  *rot_out = rot;
}

// phi slope backward (k_plm_p), gathered: cell c receives the weight on it
// of its own slope and of its two neighbours' slopes
__device__ __forceinline__ void plm_p_cell(const DevMesh &m, const AnnInfo &A, double plm, int j, int q,
                                           double *wL, double *wc, double *wR) {
  const int N = A.N;
  const long long c = A.zbase + j;
  const long long cL = A.zbase + (j == 0 ? N - 1 : j - 1);
  const long long cR = A.zbase + (j == N - 1 ? 0 : j + 1);
  const double dpL = .5 * (m.dphi[c] + m.dphi[cL]);
  const double dpR = .5 * (m.dphi[c] + m.dphi[cR]);
  plm_p_weights(m.prim[q * m.Ntot + cL], m.prim[q * m.Ntot + c], m.prim[q * m.Ntot + cR], dpL, dpR, plm, wL, wc, wR);
}
__device__ void reduce_block(const AdjExtra &x, const AdjMesh &a, int rb, int n);
// The face-slope backward and the phi-slope backward in one launch, over
// tiles of TILE - 2 cells j0 .. j0+n-1.  Thread t takes cell j0 - 1 + t (mod
// N) for t = 0 .. n+1: the face-slope gather (slope_cell) completes that
// cell's phi-slope adjoint agp, and the cell's phi limiter (k_plm_p) hands
// it on: wL agp to the left neighbour's primitives, wc agp to its own, wR
// agp to the right neighbour's; the outer two wait in LDS.  After a barrier
// each of the n cells adds its own share and its neighbours' (the two halo
// cells, their own tiles' cells, only hand theirs on).  Each cell's limiter
// weights are evaluated once, and the phi-slope adjoint never goes back to
// HBM (nothing after this substep reads it).  Workgroups past the nmain
// tiles: the reduction's blocks (whole workgroups each).
#ifndef ADJ_SLOPE_WAVES
#define ADJ_SLOPE_WAVES 1
#endif
__global__ __launch_bounds__(TILE, ADJ_SLOPE_WAVES) void ka_slope_p(DevMesh m, AdjMesh a, const int2 *tiles,
                                                                    int nmain, int nslots, AdjExtra x) {
  const int b = xcd_order();
  if (b >= nmain) {
    reduce_block(x, a, b - nmain, nslots);
    return;
  }
  __shared__ double s_toL[5][TILE], s_toR[5][TILE];
  const int2 tl = tiles[b];
  const AnnInfo A = m.ann[tl.x];
  const DiskParams &P = disk_of(m, A.disk);
  const int N = A.N;
  const int n = min(TILE - 2, N - tl.y);
  const int t = threadIdx.x;
  int j = tl.y - 1 + t;
  if (j < 0) j += N;
  if (j >= N) j -= N;
  const long long z = A.zbase + j;
  double ap[5], own[5], rot = 0.0;
  if (t <= n + 1) {
    double agp[5];
    slope_cell(m, a, A, P, j, z, ap, agp, &rot);
    for (int q = 0; q < 5; ++q) {
      const double g = a.agradp[q * m.Ntot + z] + agp[q];
      double wL, wc, wR;
      plm_p_cell(m, A, P.plm, j, q, &wL, &wc, &wR);
      own[q] = wc * g;
      s_toL[q][t] = wL * g;
      s_toR[q][t] = wR * g;
    }
  }
  __syncthreads();
  if (t < 1 || t > n) return;
  for (int q = 0; q < 5; ++q) {
    double s = own[q];
    s += s_toR[q][t - 1];  // the left neighbour's right weight
    s += s_toL[q][t + 1];  // the right neighbour's left weight
    double v = a.aP[q * m.Ntot + z] + ap[q];
    v += s;
    a.aP[q * m.Ntot + z] = v;
  }
  a.arot[z] += rot;
}

// CFL step backward.  The step's dt is the minimum of the interior cells'
// steps; on a smooth disk whole annuli can tie bit for bit, and the minimum
// is then the step of any one of them: the first tied cell (lowest zone)
// takes lam_dt times its gradient.  pick_cell finds it from the step's tile
// minima (one workgroup of the step's first launch), ka_chain applies it.  m.sc holds the scalars the forward's CFL minimum
// saw.
__device__ __forceinline__ double cfl_cell(const DevMesh &m, const int2 tl, const AnnInfo &A, const DiskParams &P,
                                           const DiskScalars &sc, long long c, double *pr, double *r_out,
                                           double *dx_out, double *w_out, double *tc_out) {
  const double rm = A.rm, rp = A.rp, dz = A.zp - A.zm, dr = rp - rm, r = .5 * (rp + rm);
  const double wj = m.wiph[tl.x];
  const double w = .5 * (wj + wj);
  double dx = dr;
  const double rdphi = .5 * (rp + rm) * m.dphi[c];
  if (rdphi < dr) dx = rdphi;
  if (dx > dz) dx = dz;
  load5(m.prim, m.Ntot, c, pr);
  const double tc = (P.visc > 0.0 && P.visc_const != 1) ? clean_pi(m.tiph[c]) - 0.5 * m.dphi[c] : 0.0;
  *r_out = r;
  *dx_out = dx;
  *w_out = w;
  *tc_out = tc;
  return cell_dt_value(P, sc, pr, r, dx, w, tc);
}
// Block rb of the reduction: buffer rb / ADJ_RED_BLOCKS of the two
// substeps' partials (x.red_src: the second substep's, then the first's),
// its share of the n slots, into x.red.
__device__ void reduce_block(const AdjExtra &x, const AdjMesh &a, int rb, int n) {
  const int buf = rb / ADJ_RED_BLOCKS, blk = rb % ADJ_RED_BLOCKS;
  const double *part = x.red_src + (long long)buf * SA_FIELDS * a.npart;
  const int per = (n + ADJ_RED_BLOCKS - 1) / ADJ_RED_BLOCKS;
  const int i0 = blk * per, i1 = min(n, i0 + per);
  double v[SA_FIELDS];
#pragma unroll
  for (int k = 0; k < SA_FIELDS; ++k) v[k] = 0.0;
#pragma unroll 1
  for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x)
#pragma unroll
    for (int k = 0; k < SA_FIELDS; ++k) v[k] += part[k * a.npart + i];
  block_sum9(v, x.red + (long long)rb * SA_FIELDS, 1);
}

// per annulus sum of the cells' rotation adjoints src into out[ann]; clears them
__device__ void rot_annulus(const DevMesh &m, const double *src0, int ann, double *out) {
  __shared__ double sh[TILE / 64];
  double *src = (double *)src0;
  const AnnInfo A = m.ann[ann];
  double s = 0.0;
  for (int j = threadIdx.x; j < A.N; j += TILE) {
    s += src[A.zbase + j];
    src[A.zbase + j] = 0.0;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < TILE / 64; ++w) t += sh[w];
    out[ann] = t;
  }
}

__global__ void ka_copy_scalars(DiskScalars *dst, const DiskScalars *src) {
  const int n = sizeof(DiskScalars) / sizeof(double);
  if ((int)threadIdx.x < n) ((double *)dst)[threadIdx.x] = ((const double *)src)[threadIdx.x];
}

// adjoints of a mass's phase and radius through xbh = r cos(phi), ybh = r sin(phi)
__device__ __forceinline__ void chain_pos(const DiskScalars &sc, const ScalarAdj &s, double *lphi, double *lr) {
  for (int p = 0; p < 2; ++p) {
    lphi[p] = -sc.ybh[p] * s.xbh[p] + sc.xbh[p] * s.ybh[p];
    const double r = sc.gm[p].r;
    lr[p] = r != 0.0 ? (sc.xbh[p] * s.xbh[p] + sc.ybh[p] * s.ybh[p]) / r : 0.0;
  }
}

static __device__ __forceinline__ ScalarAdj scalar_adj_of(const double *v) {
  ScalarAdj s;
  s.dt = v[0];
  s.alpha = v[1];
  s.M[0] = v[2];
  s.M[1] = v[3];
  s.xbh[0] = v[4];
  s.xbh[1] = v[5];
  s.ybh[0] = v[6];
  s.ybh[1] = v[7];
  s.loss = v[8];
  return s;
}

// The step's CFL cell, one workgroup: the lowest interior zone whose step
// is the minimum *dtp (the first tile, in zone order, whose minimum it is,
// then the lowest such cell in it; a tile whose cells missed it, were one
// to, would pass to the next), and the gradient of its step (one thread),
// into *pick.  m: the step-start state, the scalars the forward's CFL
// minimum saw, m.dt_tile its tile minima.
__device__ void pick_cell(const DevMesh &m, const int2 *itiles, int nitiles, int nann,
                          const unsigned long long *dtp, CflGrad *pick) {
  __shared__ int s_i[TILE / 64];
  const unsigned long long dtbits = *dtp;
  auto block_min = [&](int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_down(v, off));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_i[threadIdx.x >> 6] = v;
    __syncthreads();
    int r = s_i[0];
    for (int w = 1; w < TILE / 64; ++w) r = min(r, s_i[w]);
    return r;
  };
  int c = 0x7fffffff;
  int from = 0;
  while (c == 0x7fffffff && from < nitiles) {
    int ft = 0x7fffffff;
    for (int t = from + threadIdx.x; t < nitiles; t += TILE)
      if ((unsigned long long)__double_as_longlong(m.dt_tile[t]) == dtbits) {
        ft = t;
        break;
      }
    const int tile = block_min(ft);
    if (tile == 0x7fffffff) break;
    const int2 tl = itiles[tile];
    const AnnInfo At = m.ann[tl.x];
    const int j = tl.y + threadIdx.x;
    int cc = 0x7fffffff;
    if (j < At.N) {
      const long long z = At.zbase + j;
      double pr[5], r, dx, w, tc;
      const double d = cfl_cell(m, tl, At, disk_of(m, At.disk), m.sc[At.disk], z, pr, &r, &dx, &w, &tc);
      if ((unsigned long long)__double_as_longlong(d) == dtbits) cc = (int)z;
    }
    c = block_min(cc);
    from = tile + 1;
  }
  if (threadIdx.x != 0) return;
  CflGrad g = {};
  g.c = c;
  if (c >= 0 && c < m.Ntot) {
    int l = 0, hi = nann - 1;  // the annulus holding cell c (zbase ascending)
    while (l < hi) {
      const int mid = (l + hi + 1) >> 1;
      if (m.ann[mid].zbase <= c)
        l = mid;
      else
        hi = mid - 1;
    }
    g.ann = l;
    const AnnInfo An = m.ann[l];
    const DiskParams &P = disk_of(m, An.disk);
    double pr[5], r, dx, w, tc;
    cfl_cell(m, make_int2(l, 0), An, P, m.sc[An.disk], c, pr, &r, &dx, &w, &tc);
    ScalarAdj sb = {};
    vjp_cell_dt(P, m.sc[An.disk], pr, r, dx, w, tc, 1.0, g.prb, sb, &g.tcb);
    const double v[9] = {sb.dt, sb.alpha, sb.M[0], sb.M[1], sb.xbh[0], sb.xbh[1], sb.ybh[0], sb.ybh[1], sb.loss};
    for (int k = 0; k < 9; ++k) g.sa[k] = v[k];
  }
  *pick = g;
}

// The scalar chain of one step (one workgroup).  The two substeps' scalar
// adjoints S2, S1 are the sums of the reduction's partials.  The step's dt
// enters: both substeps' dt_s (dt, dt/2), the phase the binary advances
// (omega dt) and every annulus's turn (Omega_a dt, Omega_a = w_a / r_a,
// cell_misc.c:73-75), and T (lamT).  lam_dt = 0 when dt was clipped (the CFL
// minimum did not set it).  Then the CFL step's own backward: lam_dt times
// the gradient of the step of the cell pick_cell chose (*pick, with that
// gradient), into aP, the rotation adjoint of its annulus and its scalar
// adjoints S0.  The step's contributions to the sums: Ltheta += th1n + th2
// (+ the CFL cell's), and the masses' phases, radii and masses and alpha in
// cs.  th1n: the first substep's rotation sums of the step after this one
// (summed one step late; Ltheta is read only by the chains of earlier
// steps, so each of them still sees every later step's).  m.sc: the
// scalars the forward's CFL minimum saw (pre).
__global__ __launch_bounds__(TILE) void ka_chain(DevMesh m, AdjMesh a, const CflGrad *pick, int nann,
                                                 const DiskScalars *pre,
                                                 const DiskScalars *A, const DiskScalars *B, const double *red,
                                                 const double *th1n, const double *th2, double *Ltheta, double lamT,
                                                 AdjChain *cs) {
  __shared__ double sh[TILE / 64];
  __shared__ double s_sa[2][SA_FIELDS];  // S2, S1
  if (threadIdx.x < 2 * SA_FIELDS) {
    const int buf = threadIdx.x / SA_FIELDS, k = threadIdx.x % SA_FIELDS;
    double t = 0.0;
    for (int b = 0; b < ADJ_RED_BLOCKS; ++b) t += red[(buf * ADJ_RED_BLOCKS + b) * SA_FIELDS + k];
    s_sa[buf][k] = t;
  }
  double s = 0.0;
  for (int k = threadIdx.x; k < nann; k += TILE) {
    const AnnInfo &an = m.ann[k];
    const double v = Ltheta[k] + th1n[k] + th2[k];
    s += m.wiph[k] / (.5 * (an.rp + an.rm)) * v;
    Ltheta[k] = v;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();  // sh and s_sa
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < TILE / 64; ++w) t += sh[w];
    const ScalarAdj S2 = scalar_adj_of(s_sa[0]), S1 = scalar_adj_of(s_sa[1]);
    double lphi2[2], lr2[2];
    chain_pos(*B, S2, lphi2, lr2);
    const double omega = A->gm[0].omega;
    const double lam = 1.0 * S1.dt + 0.5 * S2.dt + omega * (cs->Lphi[0] + cs->Lphi[1] + lphi2[0] + lphi2[1]) +
                       lamT + t;
    const double dtmin = __longlong_as_double((long long)A->dtmin_bits);
    const double lam_dt = A->dt == dtmin ? lam : 0.0;
    cs->lam_dt = lam_dt;
    ScalarAdj S0 = {};
    const CflGrad &g = *pick;
    if (g.c >= 0 && g.c < m.Ntot) {  // lam_dt times the gradient of the CFL cell's step
      double prb[5];
      for (int q = 0; q < 5; ++q) prb[q] = lam_dt * g.prb[q];
      add5(a.aP, m.Ntot, g.c, prb);
      Ltheta[g.ann] += lam_dt * g.tcb;  // every thread's Ltheta writes precede the barrier above
      double v[9];
      for (int k = 0; k < 9; ++k) v[k] = lam_dt * g.sa[k];
      S0 = scalar_adj_of(v);
    }
    double lphi0[2], lr0[2], lphi1[2], lr1[2];
    chain_pos(*pre, S0, lphi0, lr0);
    chain_pos(*A, S1, lphi1, lr1);
    for (int p = 0; p < 2; ++p) {
      cs->Lphi[p] += lphi1[p] + lphi2[p] + lphi0[p];
      cs->Lr[p] += lr1[p] + lr2[p] + lr0[p];
      cs->LM[p] += S1.M[p] + S2.M[p] + S0.M[p];
    }
    cs->Lalpha += S1.alpha + S2.alpha + S0.alpha;
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
void launch_adjoint_substep(const DevMesh &m, const AdjMesh &a, const LaunchPlan &L, hipStream_t st,
                            const AdjExtra &x) {
  const dim3 B(TILE);
  // part slots: ka_update [0, nz), the phi faces [nz, nz+nu), the r faces
  // [nz+nu, nz+nu+nrf) (Riemann part), and with viscosity the same two
  // ranges of the viscous part after them
  const int nz = L.n_ztiles, nu = L.n_utiles, nrf = L.n_rftiles;
  const int nslots = a.visc ? nz + 2 * nu + 2 * nrf : nz + nu + nrf;
  const int nred = x.red ? 2 * ADJ_RED_BLOCKS : 0;
  // the rotation sums and the CFL cell ride in the update launch
  const int nrot = x.rot_src ? L.nann : 0;
  const int npick = x.pick ? 1 : 0;
  hipLaunchKernelGGL(ka_update, dim3(nz + nrot + npick), B, 0, st, m, a, L.ztiles, nz, nrot, L.itiles, L.n_itiles,
                     L.nann, x);
  hipLaunchKernelGGL(ka_faces_h, dim3(nu + nrf), B, 0, st, m, a, L.utiles, nu, L.rftiles, nz);
  if (a.visc) hipLaunchKernelGGL(ka_faces_v, dim3(nu + nrf), B, 0, st, m, a, L.utiles, nu, L.rftiles, nz + nu + nrf);
  hipLaunchKernelGGL(ka_plm_r, dim3(nz), B, 0, st, m, a, L.ztiles);
  hipLaunchKernelGGL(ka_slope_p, dim3(L.n_stiles + nred), B, 0, st, m, a, L.stiles, L.n_stiles, nslots, x);
}

void launch_copy_scalars(DiskScalars *dst, const DiskScalars *src, hipStream_t st) {
  static_assert(sizeof(DiskScalars) % sizeof(double) == 0 && sizeof(DiskScalars) <= 64 * sizeof(double),
                "one wave copies the scalars");
  hipLaunchKernelGGL(ka_copy_scalars, dim3(1), dim3(64), 0, st, dst, src);
}

void launch_adjoint_chain(const DevMesh &m, const AdjMesh &a, const CflGrad *pick, int nann, const DiskScalars *pre,
                          const DiskScalars *A, const DiskScalars *B, const double *red, const double *th1n,
                          const double *th2, double *Ltheta, double lamT, AdjChain *cs, hipStream_t st) {
  hipLaunchKernelGGL(ka_chain, dim3(1), dim3(TILE), 0, st, m, a, pick, nann, pre, A, B, red, th1n, th2, Ltheta, lamT,
                     cs);
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free torque loss (disco_batch_torque_loss_noalloc):
// slot 0 is the loss accumulator, slots 1.. the per-tile torque sums that
// k_update writes.  One copy per GPU, held until the process ends; every
// call on a GPU shares it, so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[TORQUE_SCRATCH_DOUBLES];

__global__ __launch_bounds__(64) void ka_zero(double *p, int n) {
  for (int i = threadIdx.x; i < n; i += 64) p[i] = 0.0;
}
void launch_zero(double *p, int n, hipStream_t st) { hipLaunchKernelGGL(ka_zero, dim3(1), dim3(64), 0, st, p, n); }

// The device address of this GPU's scratch (the array's name is only a
// host-side stand-in): looked up once per device, with that device current.
double *torque_scratch_device() {
  static std::mutex mu;
  static double *addr[TORQUE_SCRATCH_MAX_DEVICES] = {};
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= TORQUE_SCRATCH_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    addr[dev] = (double *)p;
  }
  return addr[dev];
}
