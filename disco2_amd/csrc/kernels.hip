// HIP kernels of the hydro substep for gfx950 (FP64).
//
// Every per-zone kernel runs on "zone tiles": one workgroup owns up to
// DISCO_TILE consecutive cells of one annulus, so loads along phi are
// coalesced and the annulus constants are uniform in the workgroup.
// Per-cell sums are gathers in the reference's accumulation order (never
// atomics), which keeps the results deterministic and bitwise comparable.
#include <hip/hip_runtime.h>

#include "cell_ops.h"
#include "device_common.h"

// ----------------------------------------------------------------------------
// Overlap faces between two adjacent annuli: build_jloop and face_create,
// src/Face/face_create_destroy.c:22-193.
//
// The reference walks the inner annulus sequentially, carrying the index jp
// of the outer cell that contains the current inner cell's left edge.  Here
// every inner cell j guesses its starting jp (the outer cell containing its
// left edge), runs the reference's per-cell step from that guess, and the
// guesses are verified against the previous cell's final jp.  When every
// guess matches, the result equals the sequential walk by induction; when
// one does not (ties between edges), one thread re-runs the sequential walk.
// Faces are stored per pair: L/R zone, width dphi and centre cm.  A face of
// inner cell j that advances jp starts a new "run" of the outer cell; run t
// of the pair begins at run_start[t], and the outer cell c owns runs
// t = (c - jpmin) mod N_out (+ N_out ...), which gives the R-side gathers
// the reference's accumulation order.
// ----------------------------------------------------------------------------
struct FaceOut {
  int *fL, *fR, *run_start;
  signed char *ftype;  // optional: which edges bound the face (FACE_II, _IO, _OO, _OI)
  double *fdp, *fcm;
  long long fbase;
  int rbase, cap, run_cap;
  long long zin, zout;
  int *err;
};

template <bool EMIT>
__device__ int jloop_cell(int j, int jps, const double *Tin, const double *Din, const double *Tout,
                          const double *Dout, int Nin, int Nout, int foff, const FaceOut &o, int *jpe,
                          int *bad) {
  int jm = j - 1;
  if (jm < 0) jm = Nin - 1;
  int jp = jps;
  int n = 0;
  auto put = [&](int R, double dph, double tp, int type) {
    if (EMIT) {
      int f = foff + n;
      if (f < o.cap) {
        long long g = o.fbase + f;
        o.fL[g] = (int)(o.zin + j);
        o.fR[g] = (int)(o.zout + R);
        o.fdp[g] = dph;
        o.fcm[g] = tp - .5 * dph;  // addFace, :11-20
        if (o.ftype) o.ftype[g] = (signed char)type;
      } else {
        set_err(o.err, DERR_FACE_OVERFLOW);
      }
      if (n >= 1) {
        int t = foff - j + n;
        if (t < o.run_cap)
          o.run_start[o.rbase + t] = f;
        else
          set_err(o.err, DERR_FACE_OVERFLOW);
      }
    }
    ++n;
  };
  double dp = wrap_p(Tout[jp] - Tin[jm]);
  if (Din[j] < dp) {
    put(jp, Din[j], Tin[j], FACE_II);
  } else {
    put(jp, dp, Tout[jp], FACE_IO);  // step A
    ++jp;
    if (jp == Nout) jp = 0;
    dp = wrap_pm(Tin[j] - Tout[jp]);
    int guard = 0;
    while (dp > 0.0) {  // step B
      if (++guard > Nout) {
        *bad = 1;
        break;
      }
      put(jp, Dout[jp], Tout[jp], FACE_OO);
      ++jp;
      if (jp == Nout) jp = 0;
      dp = wrap_pm(Tin[j] - Tout[jp]);
    }
    dp = Dout[jp] + dp;
    put(jp, dp, Tin[j], FACE_OI);  // step C
  }
  *jpe = jp;
  return n;
}

// outer cell whose right edge is the first one at or after angle e
__device__ int guess_outer(double e, const double *Tout, int Nout) {
  double x = wrap_p(e - Tout[Nout - 1]);
  if (!(x >= 0.0)) x = 0.0;
  int c = (int)(x * (double)Nout / TWO_PI);
  if (c < 0) c = 0;
  if (c >= Nout) c = Nout - 1;
  for (int it = 0; it <= Nout && wrap_pm(e - Tout[c]) > 0.0; ++it) c = (c + 1 == Nout) ? 0 : c + 1;
  for (int it = 0; it <= Nout; ++it) {
    int cm = (c == 0) ? Nout - 1 : c - 1;
    if (wrap_pm(e - Tout[cm]) <= 0.0)
      c = cm;
    else
      break;
  }
  return c;
}

// threads of the face builder (one workgroup per annulus pair): 1024, 512
// or 256, a launch-schedule choice (the faces do not depend on it)

// exclusive prefix sum over the workgroup (wave shuffles, then wave totals)
__device__ int block_excl_scan(int v, int *s_w, int *total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_w[w] = x;
  __syncthreads();
  if (w == 0) {
    int t = lane < nw ? s_w[lane] : 0;
    for (int o = 1; o < 64; o <<= 1) {
      int y = __shfl_up(t, o);
      if (lane >= o) t += y;
    }
    if (lane < nw) s_w[lane] = t;
  }
  __syncthreads();
  int base = w > 0 ? s_w[w - 1] : 0;
  *total = s_w[nw - 1];
  __syncthreads();
  return base + x - v;
}

// One workgroup of FTB threads per annulus pair builds the pair's faces.
// Inner cell j's walk starts at the outer cell holding its left edge and
// ends at the one holding its right edge, emitting one face more than the
// cursor advances.  So with g(j) = guess_outer(Tin[j]) (and g(-1) = jpmin,
// face_create_destroy.c:82-96) the face count of cell j is
// 1 + (g(j) - g(j-1)) mod N_out.  Cells are taken FTB at a time: counts,
// workgroup prefix sum, then every cell walks from g(j-1) with the
// reference's per-cell step writing its faces, and checks that it wrote
// that many faces and ended at g(j).  If all cells agree, every walk began
// where the sequential walk would have (induction over j) and the faces are
// the reference's; if any did not (ties between edges), thread 0 replays the
// sequential walk.  The outer annulus's edges are staged in LDS when they
// fit, so the walks' dependent loads are LDS reads.
//
// A pair of equal annuli whose cells all have two faces is regular
// (PairState::reg): its faces follow from the edges alone, and the step's
// kernels compute them where they use them.  With skip set, such a pair's
// walks are verified without storing its faces, face offsets and runs;
// the class API's face export, the adjoint's record and the outflow
// boundary's pairs (r pairs from annuli 0, NR-3, NR-2, whose remap reads
// the stored faces) always store them.
template <int FTB>
__device__ __forceinline__ void faces_pair(const DevMesh &m, const PairInfo &P, const AnnInfo &A, const AnnInfo &B,
                                           int p, const double *Tout, const double *Dout, double *s_d, int *s_j,
                                           int *s_w, int skip) {
  const int tid = threadIdx.x;
  const int Nin = A.N, Nout = B.N;
  const double *Tin = m.tiph + A.zbase, *Din = m.dphi + A.zbase;
  int *fo = (P.dir == 0 ? m.fo_r : m.fo_z) + A.zbase;
  // jpmin: smallest wrap_p(Tout[jp] - Tin[Nin-1]), first index on ties
  {
    const double p0 = Tin[Nin - 1];
    double bd = TWO_PI;
    int bj = -1;
    for (int jp = tid; jp < Nout; jp += FTB) {
      double dp = wrap_p(Tout[jp] - p0);
      if (bd > dp) {
        bd = dp;
        bj = jp;
      }
    }
    s_d[tid] = bd;
    s_j[tid] = bj;
    __syncthreads();
    for (int s = FTB / 2; s > 0; s >>= 1) {
      if (tid < s) {
        int j2 = s_j[tid + s];
        if (j2 >= 0) {
          int j1 = s_j[tid];
          double d1 = s_d[tid], d2 = s_d[tid + s];
          if (j1 < 0 || d2 < d1 || (d2 == d1 && j2 < j1)) {
            s_d[tid] = d2;
            s_j[tid] = j2;
          }
        }
      }
      __syncthreads();
    }
  }
  const int jpmin = s_j[0] < 0 ? 0 : s_j[0];
  if (skip && Nin == Nout && 2 * Nin <= P.cap && Nin + 1 < P.run_cap) {
    const int NR = disk_of(m, P.disk).NR;
    if (P.dir == 1 || (A.i != 0 && A.i < NR - 3)) {
      // every cell two faces, each walk as guessed: regular, nothing stored
      int irr = 0;
      for (int base = 0; base < Nin; base += FTB) {
        const int j = base + tid;
        if (j < Nin) {
          const int gs = (j == 0) ? jpmin : guess_outer(Tin[j - 1], Tout, Nout);
          const int ge = guess_outer(Tin[j], Tout, Nout);
          int adv = ge - gs;
          if (adv < 0) adv += Nout;
          int jpe, b2 = 0;
          FaceOut none{};
          const int n = adv == 1 ? jloop_cell<false>(j, gs, Tin, Din, Tout, Dout, Nin, Nout, 2 * j, none, &jpe, &b2)
                                 : 0;
          irr |= adv != 1 || b2 || n != 2 || (j + 1 < Nin && jpe != ge);
        }
      }
      irr = __syncthreads_or(irr);
      if (!irr) {
        if (tid == 0) {
          PairState st;
          st.count = 2 * Nin;
          st.jpmin = jpmin;
          st.T = Nin;
          st.seq = 0;
          st.reg = 1;
          m.pst[p] = st;
        }
        return;
      }
    }
  }
  FaceOut o;
  o.fL = m.fL;
  o.fR = m.fR;
  o.ftype = m.ftype;
  o.fdp = m.fdphi;
  o.fcm = m.fcm;
  o.run_start = m.run_start;
  o.fbase = P.face_base;
  o.rbase = P.run_base;
  o.cap = P.cap;
  o.run_cap = P.run_cap;
  o.zin = A.zbase;
  o.zout = B.zbase;
  o.err = m.err;
  int bad = 0, carry = 0, irregular = 0;
  for (int base = 0; base < Nin; base += FTB) {
    const int j = base + tid;
    int gs = 0, ge = 0, cnt = 0;
    if (j < Nin) {
      gs = (j == 0) ? jpmin : guess_outer(Tin[j - 1], Tout, Nout);
      ge = guess_outer(Tin[j], Tout, Nout);
      int adv = ge - gs;
      if (adv < 0) adv += Nout;
      cnt = 1 + adv;
      irregular |= cnt != 2;
    }
    int chunk;
    const int foff = carry + block_excl_scan(cnt, s_w, &chunk);
    if (j < Nin) {
      fo[j] = foff;
      int jpe, b2 = 0;
      int n = jloop_cell<true>(j, gs, Tin, Din, Tout, Dout, Nin, Nout, foff, o, &jpe, &b2);
      if (b2 || n != cnt || (j + 1 < Nin && jpe != ge)) bad = 1;
    }
    carry += chunk;
  }
  bad = __syncthreads_or(bad);
  irregular = __syncthreads_or(irregular);
  int total = carry;
  if (bad && tid == 0) {
    int f = 0, b2 = 0, jp = jpmin;
    for (int j = 0; j < Nin; ++j) {
      fo[j] = f;
      int jpe;
      f += jloop_cell<true>(j, jp, Tin, Din, Tout, Dout, Nin, Nout, f, o, &jpe, &b2);
      jp = jpe;
    }
    total = f;
  }
  if (tid == 0) {
    int T = total - Nin;
    if (total > P.cap || T + 1 >= P.run_cap || T < 0) {
      set_err(m.err, DERR_FACE_OVERFLOW);
      T = -1;
      total = 0;
    } else {
      m.run_start[P.run_base] = 0;
      m.run_start[P.run_base + T + 1] = total;
    }
    PairState st;
    st.count = total;
    st.jpmin = jpmin;
    st.T = T;
    st.seq = bad;
    // every cell verified with two faces: the walk advanced one outer cell
    // per inner cell from jpmin
    st.reg = !bad && !irregular && Nin == Nout && total == 2 * Nin;
    m.pst[p] = st;
  }
}

// nstage: LDS doubles per array for the outer annulus (0: none staged)
// Two workgroups per CU (all 512 annulus pairs of a 512-annulus disk
// resident at once) need 8 waves per SIMD: left alone the compiler takes 106
// SGPRs, which allows 7, so one workgroup per CU and two passes over the
// CUs.
template <int FTB>
__global__ __launch_bounds__(FTB) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_faces(DevMesh m, int pair0,
                                                                                           int nstage, int skip) {
  extern __shared__ double s_stage[];  // Tout [nstage], Dout [nstage]
  __shared__ double s_d[FTB];
  __shared__ int s_j[FTB];
  __shared__ int s_w[FTB / 64];
  const int b = pair0 + tile_of(m);
  const int p = m.pair_perm ? m.pair_perm[b] : b;
  const PairInfo P = m.pairs[p];
  const AnnInfo A = m.ann[P.ann_in];
  const AnnInfo B = m.ann[P.ann_out];
  const double *Tg = m.tiph + B.zbase, *Dg = m.dphi + B.zbase;
  if (B.N <= nstage) {
    for (int i = threadIdx.x; i < B.N; i += FTB) {
      s_stage[i] = Tg[i];
      s_stage[nstage + i] = Dg[i];
    }
    __syncthreads();
    faces_pair<FTB>(m, P, A, B, p, s_stage, s_stage + nstage, s_d, s_j, s_w, skip);
  } else {
    faces_pair<FTB>(m, P, A, B, p, Tg, Dg, s_d, s_j, s_w, skip);
  }
}

// ----------------------------------------------------------------------------
// Mesh velocity per annulus: cell_clear_w + cell_set_w, src/Cell/cell_set_w.c:12-102.
// One wavefront per annulus; C_RIGID sums in the reference's cell order.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_set_w(DevMesh m, int ann0) {
  const int a = ann0 + tile_of(m);
  const AnnInfo A = m.ann[a];
  const DiskParams &P = disk_of(m, A.disk);
  const int lane = threadIdx.x;
  const double r = 0.5 * (A.rm + A.rp);
  double w = 0.0;
  if (P.w_numeric == C_RIGID) {
    __shared__ double s_vm[64], s_m[64];
    const double dz = A.zp - A.zm;
    const double area = .5 * (A.rp * A.rp - A.rm * A.rm);
    double ww = 0.0, Mring = 0.0;
    for (int base = 0; base < A.N; base += 64) {
      int j = base + lane;
      if (j < A.N) {
        long long z = A.zbase + j;
        double dV = area * m.dphi[z] * dz;  // cell_calc_cons volume, cell_conversion.c:45
        double mass = m.prim[RHO * m.Ntot + z] * dV;
        double vp = r * m.prim[UPP * m.Ntot + z];
        s_vm[lane] = vp * mass;
        s_m[lane] = mass;
      }
      __syncthreads();
      if (lane == 0) {
        int n = min(64, A.N - base);
        for (int x = 0; x < n; ++x) {
          Mring += s_m[x];
          ww += s_vm[x];
        }
      }
      __syncthreads();
    }
    w = ww / Mring;
  } else if (P.w_numeric == C_KEPLER) {
    w = pow(r, -0.5);
  } else if (P.w_numeric == C_OMEGA20) {
    w = r * 20.;
  } else if (P.w_numeric == C_MILOS) {
    w = (1. - exp(-pow(r, 1.5))) / sqrt(r);
  } else if (P.w_numeric == C_POWERLAW) {
    w = pow(r, -1.0);
  }
  if (lane == 0) m.wiph[a] = w;
}

// ----------------------------------------------------------------------------
// CFL time step over interior cells, cell_mindt, src/Cell/cell_mindt.c:10-100.
// Each workgroup stores the minimum of its tile with a plain store (no
// device-wide atomics: thousands of them on one word serialise); k_scalar's
// SC_DT_FOLD reduces a disk's tiles.  min is exact, so the order is free.
// ----------------------------------------------------------------------------
// the CFL step of cell c (local index j) of annulus A, tile tl, from its
// primitives pr and width dp
template <int FEAT>
__device__ __forceinline__ double cell_dt_at(const DevMesh &m, const int2 tl, const AnnInfo &A, long long c,
                                             const double *pr, double dp) {
  const DiskParams &P = disk_of(m, A.disk);
  const double rm = A.rm, rp = A.rp;
  const double dz = A.zp - A.zm;
  const double dr = rp - rm;
  const double r = .5 * (rp + rm);
  const double wj = m.wiph[tl.x];
  double w = .5 * (wj + wj);
  double dx = dr;
  double rdphi = .5 * (rp + rm) * dp;
  if (rdphi < dr) dx = rdphi;
  if (dx > dz) dx = dz;
  const double tc = ((FEAT & F_VISC) && P.visc > 0.0 && P.visc_const != 1) ? clean_pi(m.tiph[c]) - 0.5 * dp : 0.0;
  return cell_dt_value<FEAT>(P, m.sc[A.disk], pr, r, dx, w, tc);
}
template <int NQ, int FEAT = F_ALL>
__device__ __forceinline__ double cell_dt(const DevMesh &m, const int2 tl, const AnnInfo &A) {
  const int j = tl.y + threadIdx.x;
  double dt = 1.e100;
  if (j < A.N) {
    const long long c = A.zbase + j;
    double pr[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) pr[q] = m.prim[q * m.Ntot + c];
    double d = cell_dt_at<FEAT>(m, tl, A, c, pr, m.dphi[c]);
    if (d < dt) dt = d;
  }
  return dt;
}

// the workgroup's minimum of v, stored by thread 0 at dt_tile[slot]
__device__ __forceinline__ void store_tile_min(const DevMesh &m, double v, int slot) {
  __shared__ double s_m[TILE / 64];
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_down(v, off);
    if (o < v) v = o;
  }
  if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = s_m[0];
    for (int w = 1; w < TILE / 64; ++w)
      if (s_m[w] < b) b = s_m[w];
    m.dt_tile[slot] = b;
  }
}

template <int NQ>
__global__ __launch_bounds__(TILE) void k_mindt(DevMesh m, const int2 *tiles) {
  const int it = tile_of(m);
  const int2 tl = tiles[it];
  const AnnInfo A = m.ann[tl.x];
  store_tile_min(m, cell_dt<NQ>(m, tl, A), it);
}


// ----------------------------------------------------------------------------
// phi slopes, cell_plm_p, src/Cell/cell_plm.c:116-161
// ----------------------------------------------------------------------------
// DT: the first substep's launch also takes the CFL minimum of its tile
// (k_mindt's value, from the same prim and dphi) when the tile is interior,
// into the tile's slot of the interior tile list, saving k_mindt's pass.
template <int NQ, bool DT, int FEAT = F_ALL>
__global__ __launch_bounds__(TILE) void k_plm_p(DevMesh m, const int2 *tiles) {
  const int2 tl = tiles[tile_of(m)];
  const AnnInfo A = m.ann[tl.x];
  const int j = tl.y + threadIdx.x;
  double dtc = 1.e100;  // this cell's CFL step (DT, interior tiles)
  if (j < A.N) {
    const double PLM = m.dp[A.disk].plm;
    const int N = A.N;
    const long long c = A.zbase + j;
    const long long cL = A.zbase + (j == 0 ? N - 1 : j - 1);
    const long long cR = A.zbase + (j == N - 1 ? 0 : j + 1);
    const double dpc = m.dphi[c];
    const double dpL = .5 * (dpc + m.dphi[cL]);
    const double dpR = .5 * (dpc + m.dphi[cR]);
    double pr[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) pr[q] = m.prim[q * m.Ntot + c];
    if (DT && A.itile >= 0) {
      const double d = cell_dt_at<FEAT>(m, tl, A, c, pr, dpc);
      if (d < dtc) dtc = d;  // as cell_dt: a NaN or huge step counts as 1e100
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const double *pq = m.prim + q * m.Ntot;
      double vc = pr[q], vL = pq[cL], vR = pq[cR];
      double sL = (vc - vL) / dpL;
      double sR = (vR - vc) / dpR;
      double sM = (vR - vL) / (dpL + dpR);
      double s = PLM * sL;
      if (fabs(s) > PLM * fabs(sR)) s = PLM * sR;
      if (fabs(s) > fabs(sM)) s = sM;
      if (sL * sR < 0.0) s = 0.0;
      m.gradp[q * m.Ntot + c] = s;
    }
  }
  if (DT && A.itile >= 0) store_tile_min(m, dtc, A.itile + tl.y / TILE);
}

// ----------------------------------------------------------------------------
// r or z slopes, cell_plm_rz, src/Cell/cell_plm.c:13-114, as a per-cell
// gather: the cell's faces as R (previous pair) then as L (next pair), in
// face order, for the sum and again for the limiter.
// ----------------------------------------------------------------------------
// The face difference S = (W_R - W_L)/(deltaL + deltaR) (cell_plm.c:36-49 and
// 86-99) is computed where it is used, from the face's two cells, with
// face_slope's arithmetic: the cell's own state is loaded once, the other
// cell's per face (a neighbour's cells sit in the same L2).  Each face is
// evaluated by both of its cells and never goes through HBM.  The slopes of
// a cell's first PlmCache::NC faces wait in LDS (a private column per thread, no
// barrier) between the sum and the limiter's walk; the rest, rare, are
// evaluated again.  LDS rather than registers keeps the kernel at four waves
// per SIMD.
template <int NQ>
struct PlmCache {
  static constexpr int NC = NQ <= 6 ? 4 : 3;  // 40-48 KiB per workgroup
};
// S of a face (centre cm, other cell oth) of a cell whose centre angle
// clean_pi(tiph) - dphi/2 is phi_o, primitives w and phi slopes gp;
// own_left: the cell is the face's L
template <int NQ>
__device__ __forceinline__ void own_face_slope(const DevMesh &m, double phi_o, const double *w, const double *gp,
                                               bool own_left, double dsum, long long oth, double cm, double *S) {
  const double ph = clean_pi(m.tiph[oth]) - .5 * m.dphi[oth];
  const double pL = own_left ? phi_o : ph;
  const double pR = own_left ? ph : phi_o;
  const double dpL = wrap_pm(cm - pL);
  const double dpR = wrap_pm(pR - cm);
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const double v = m.prim[q * m.Ntot + oth], gv = m.gradp[q * m.Ntot + oth];
    const double WL = own_left ? w[q] + dpL * gp[q] : v + dpL * gv;
    const double WR = own_left ? v - dpR * gv : w[q] - dpR * gp[q];
    S[q] = (WR - WL) / dsum;
  }
}
// The r (dir 0) or z (dir 1) slope of cell j (zone z) of annulus A into
// out[q * Ntot + z], from the cell's primitives w, phi slopes gp, centre angle
// phi_o and width dp; s_S: the thread's LDS column of cached face slopes.
// The face ranges of cell j (zone z) of an annulus of N cells: as R of pair
// pR, up to two runs (the second is the wrap), into b[0..1], e[0..1]; as L
// of pair pL, one range, into b[2], e[2] (faces_as_R / faces_as_L order).
// Absent pairs (-1) leave empty ranges.
__device__ __forceinline__ void cell_face_ranges(const DevMesh &m, int pR, int pL, const int *fo, long long z, int j,
                                                 int N, int *b, int *e) {
  b[0] = b[1] = b[2] = 0;
  e[0] = e[1] = e[2] = 0;
  if (pR >= 0) {
    const PairState S = m.pst[pR];
    int tt = j - S.jpmin;
    if (tt < 0) tt += N;
    if (S.reg) {  // runs of two faces 2t - 1, 2t; the wrap cell's two runs are faces 0 and 2N - 1
      if (tt == 0) {
        e[0] = 1;
        b[1] = 2 * N - 1;
        e[1] = 2 * N;
      } else {
        b[0] = 2 * tt - 1;
        e[0] = 2 * tt + 1;
      }
    } else {
      const PairInfo PI = m.pairs[pR];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        if (tt <= S.T) {
          b[u] = m.run_start[PI.run_base + tt];
          e[u] = m.run_start[PI.run_base + tt + 1];
        }
        tt += N;
      }
    }
  }
  if (pL >= 0) {
    const PairState S = m.pst[pL];
    if (S.reg) {  // faces 2j, 2j + 1
      b[2] = 2 * j;
      e[2] = 2 * j + 2;
    } else {
      int f1 = (j + 1 < N) ? fo[z + 1] : S.count;
      if (f1 > S.count) f1 = S.count;
      int f0 = fo[z];
      if (f0 < 0) f0 = 0;
      b[2] = f0;
      e[2] = f1;
    }
  }
}

// the pairs below (pR) and above (pL) of the cell in direction dir (-1: none)
__device__ __forceinline__ void plm_pairs(const AnnInfo &A, const DiskParams &P, int dir, int *pR, int *pL) {
  const int k = A.k, i = A.i;
  *pR = *pL = -1;
  if (dir == 0) {
    if (i >= 1) *pR = rpair_of(P, k, i - 1);
    if (i <= P.NR - 2) *pL = rpair_of(P, k, i);
  } else {
    if (k >= 1) *pR = zpair_of(P, k - 1, i);
    if (k <= P.NZ - 2) *pL = zpair_of(P, k, i);
  }
}

// b, e: the cell's face ranges in direction dir (cell_face_ranges), when the
// caller has loaded them already
template <int NQ, int NC = PlmCache<NQ>::NC>
__device__ __forceinline__ void plm_cell_dir(const DevMesh &m, const AnnInfo &A, const DiskParams &P, int j,
                                             long long z, int dir, const double *w, const double *gp, double phi_o,
                                             double dp, double (*s_S)[NQ][TILE], double *out, const int *b,
                                             const int *e) {
  const int i = A.i;
  double g[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) g[q] = 0.0;
  int pR, pLf;
  plm_pairs(A, P, dir, &pR, &pLf);
  // The cell's faces in the reference's order: as R, its runs in the pair
  // below (faces_as_R: at most two, the second is the wrap), then as L one
  // range in the pair above (faces_as_L); range u is faces [fb[u] + b[u],
  // fb[u] + e[u]) with its pair's r, deltaPerp and deltaL + deltaR.
  // A regular pair's faces (PairState::reg) are computed from the two
  // annuli's edges instead of loaded: rg[u], with its jpmin, cell count and
  // the inner and outer annuli's first zones.
  long long fb[3] = {0, 0, 0}, zi[3] = {0, 0, 0}, zo[3] = {0, 0, 0};
  double pr[3] = {0, 0, 0}, pp[3] = {0, 0, 0}, ds[3] = {1, 1, 1};
  int rg[3] = {0, 0, 0}, jpm[3] = {0, 0, 0}, np[3] = {0, 0, 0};
  auto pair_consts = [&](int p, int u) {
    const PairInfo PI = m.pairs[p];
    const PairState S = m.pst[p];
    fb[u] = PI.face_base;
    pr[u] = PI.r;
    pp[u] = PI.deltaPerp;
    ds[u] = PI.deltaR + PI.deltaL;
    rg[u] = S.reg;
    if (S.reg) {
      const AnnInfo Ai = m.ann[PI.ann_in];
      jpm[u] = S.jpmin;
      np[u] = Ai.N;
      zi[u] = Ai.zbase;
      zo[u] = m.ann[PI.ann_out].zbase;
    }
  };
  if (pR >= 0) {
    pair_consts(pR, 0);
    pair_consts(pR, 1);
  }
  if (pLf >= 0) pair_consts(pLf, 2);
  auto geo = [&](int u, long long gi) {
    if (rg[u]) return pair_face(m, jpm[u], np[u], zi[u], zo[u], (int)(gi - fb[u]));
    FaceGeo G;
    G.cL = m.fL[gi];
    G.cR = m.fR[gi];
    G.w = m.fdphi[gi];
    G.cm = m.fcm[gi];
    return G;
  };
  const int t = threadIdx.x;
  int n = 0;
  // the slope sum in face order; two faces per trip, so that the second
  // face's loads are in flight with the first's
  auto add_face = [&](const double *S, double dA) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) g[q] += S[q] * dA;
    if (n < NC)
#pragma unroll
      for (int q = 0; q < NQ; ++q) s_S[n][q][t] = S[q];
    ++n;
  };
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    int f = b[u];
    for (; f + 1 < e[u]; f += 2) {
      const long long gi = fb[u] + f;
      const FaceGeo G0 = geo(u, gi), G1 = geo(u, gi + 1);
      double S0[NQ], S1[NQ];
      own_face_slope<NQ>(m, phi_o, w, gp, u == 2, ds[u], u == 2 ? G0.cR : G0.cL, G0.cm, S0);
      own_face_slope<NQ>(m, phi_o, w, gp, u == 2, ds[u], u == 2 ? G1.cR : G1.cL, G1.cm, S1);
      const double dA0 = pr[u] * G0.w * pp[u];
      const double dA1 = pr[u] * G1.w * pp[u];
      add_face(S0, dA0);
      add_face(S1, dA1);
    }
    if (f < e[u]) {
      const FaceGeo G = geo(u, fb[u] + f);
      double S[NQ];
      own_face_slope<NQ>(m, phi_o, w, gp, u == 2, ds[u], u == 2 ? G.cR : G.cL, G.cm, S);
      add_face(S, pr[u] * G.w * pp[u]);
    }
  }
  const double dz = A.zp - A.zm;
  const double rp = A.rp, rm = A.rm;
  double dAtot;
  if (dir == 0) {
    if (i == 0)
      dAtot = dz * rp * dp;  // inner global edge (single rank)
    else if (i == P.NR - 1)
      dAtot = dz * rm * dp;
    else
      dAtot = dz * (rp + rm) * dp;
  } else {
    dAtot = 2. * .5 * (rp * rp - rm * rm) * dp;
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) g[q] /= dAtot;
  const double PLM = P.plm;
  n = 0;
#pragma unroll
  for (int u = 0; u < 3; ++u)
    for (int f = b[u]; f < e[u]; ++f) {
      double S[NQ];
      if (n < NC) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) S[q] = s_S[n][q][t];
      } else {
        const FaceGeo G = geo(u, fb[u] + f);
        own_face_slope<NQ>(m, phi_o, w, gp, u == 2, ds[u], u == 2 ? G.cR : G.cL, G.cm, S);
      }
      ++n;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        if (S[q] * g[q] < 0.0)
          g[q] = 0.0;
        else if (fabs(PLM * S[q]) < fabs(g[q]))
          g[q] = PLM * S[q];
      }
    }
#pragma unroll
  for (int q = 0; q < NQ; ++q) out[q * m.Ntot + z] = g[q];
}

// the cell's primitives, phi slopes, centre angle and width
template <int NQ>
__device__ __forceinline__ void plm_own(const DevMesh &m, long long z, double *w, double *gp, double *phi_o,
                                        double *dp) {
  *dp = m.dphi[z];
  *phi_o = clean_pi(m.tiph[z]) - .5 * *dp;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    w[q] = m.prim[q * m.Ntot + z];
    gp[q] = m.gradp[q * m.Ntot + z];
  }
}

template <int NQ>
__global__ __launch_bounds__(TILE) void k_plm_rz(DevMesh m, const int2 *tiles, int dir) {
  const int2 tl = tiles[tile_of(m)];
  const AnnInfo A = m.ann[tl.x];
  const int j = tl.y + threadIdx.x;
  if (j >= A.N) return;
  const long long z = A.zbase + j;
  double w[NQ], gp[NQ], phi_o, dp;
  plm_own<NQ>(m, z, w, gp, &phi_o, &dp);
  __shared__ double s_S[PlmCache<NQ>::NC][NQ][TILE];
  const DiskParams &P = disk_of(m, A.disk);
  int pR, pL, b[3], e[3];
  plm_pairs(A, P, dir, &pR, &pL);
  cell_face_ranges(m, pR, pL, dir == 0 ? m.fo_r : m.fo_z, z, j, A.N, b, e);
  plm_cell_dir<NQ>(m, A, P, j, z, dir, w, gp, phi_o, dp, s_S, m.grad, b, e);
}

// Both slopes of a 3D disk's cells in one pass (the cell's own state loaded
// once): the r slopes into grad_r, then the z slopes into m.grad, which then
// holds what the reference's grad holds after its z pass.
// face slopes of a cell cached in LDS and minimum waves per SIMD of the
// two-direction kernel
#ifndef PLM2_NC
#define PLM2_NC PlmCache<NQ>::NC
#endif
#ifndef PLM2_WAVES
#define PLM2_WAVES 1
#endif
template <int NQ>
__global__ __launch_bounds__(TILE, PLM2_WAVES) void k_plm_rz2(DevMesh m, const int2 *tiles, double *grad_r) {
  const int2 tl = tiles[tile_of(m)];
  const AnnInfo A = m.ann[tl.x];
  const int j = tl.y + threadIdx.x;
  if (j >= A.N) return;
  const long long z = A.zbase + j;
  const DiskParams &P = disk_of(m, A.disk);
  // (loading the z ranges before the r slope, to overlap their latency,
  // measured 2% slower: more registers held through the r pass)
  double w[NQ], gp[NQ], phi_o, dp;
  plm_own<NQ>(m, z, w, gp, &phi_o, &dp);
  constexpr int NC = PLM2_NC;
  __shared__ double s_S[NC][NQ][TILE];
#pragma unroll
  for (int dir = 0; dir < 2; ++dir) {
    int pR, pL, b[3], e[3];
    plm_pairs(A, P, dir, &pR, &pL);
    cell_face_ranges(m, pR, pL, dir == 0 ? m.fo_r : m.fo_z, z, j, A.N, b, e);
    plm_cell_dir<NQ, NC>(m, A, P, j, z, dir, w, gp, phi_o, dp, s_S, dir == 0 ? grad_r : m.grad, b, e);
  }
}

// ----------------------------------------------------------------------------
// Fluxes.  Face (j, j+1) of every annulus (riemann_setup_p,
// src/Riemann/Riemann_setup.c:46-80) and every r/z overlap face
// (riemann_setup_rz, :11-44), each solved once; riemann_AddFlux's
// increments dt*dA*F are stored per face and gathered by k_update.
// ----------------------------------------------------------------------------
// Face (j, j+1) of an annulus, riemann_setup_p (src/Riemann/Riemann_setup.c:46-80)
// + riemann_AddFlux: the increments X = dt*dA*F and viscous Y.
template <int NQ, int FEAT>
__device__ __forceinline__ int phi_face_flux(const DevMesh &m, const AnnInfo &A, int a, int j, const DiskParams &P,
                                             const DiskScalars &sc, double dt, double rdr_om, double *X, double *Y) {
  const long long cL = A.zbase + j;
  const long long cR = A.zbase + (j == A.N - 1 ? 0 : j + 1);
  const double dpL = m.dphi[cL], dpR = m.dphi[cR];
  const double dz = A.zp - A.zm;
  const double dr = A.rp - A.rm;
  const double r = .5 * (A.rp + A.rm);
  const double dA = dr * dz;
  double primL[NQ], primR[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    primL[q] = m.prim[q * m.Ntot + cL] + 0.5 * m.gradp[q * m.Ntot + cL] * dpL;
    primR[q] = m.prim[q * m.Ntot + cR] - 0.5 * m.gradp[q * m.Ntot + cR] * dpR;
  }
  ViscIn vi;
  if ((FEAT & F_VISC) && P.visc > 0.0) {
    visc_cells(m, vi, cL, cR);
    vi.cm = vi.tL;
    vi.r_cell_L = r;
    vi.r_cell_R = r;
    vi.rdr_om = rdr_om;
  }
  Y[0] = Y[1] = Y[2] = 0.0;
  return d_face_flux<NQ, 0, FEAT>(primL, primR, 1, r, dA, m.wiph[a], dt, P, sc, &vi, X, Y);
}

// The increments riemann_AddFlux adds for r/z overlap face g of pair PI
// (riemann_setup_rz, src/Riemann/Riemann_setup.c:11-44, and the flux of
// Riemann_routines.c:314-397): X = dt*dA*F and the viscous Y.  Stored per
// face by k_flux_rz, or evaluated inside k_update by each of the face's two
// cells (the fused schedule): the same arithmetic either way.
template <int NQ, int FEAT>
__device__ __forceinline__ int rz_face_increments(const DevMesh &m, const PairInfo &PI, const DiskParams &P,
                                                  const DiskScalars &sc, double dt, const FaceGeo &G, double *X,
                                                  double *Y) {
  const long long cL = G.cL, cR = G.cR;
  const double cm = G.cm;
  const double tL = clean_pi(m.tiph[cL]), tR = clean_pi(m.tiph[cR]);
  const double dL = m.dphi[cL], dR = m.dphi[cR];
  double pL = tL - .5 * dL;
  double pR = tR - .5 * dR;
  double dpL = cm - pL;
  double dpR = -cm + pR;
  dpL = wrap_pm(dpL);
  dpR = wrap_pm(dpR);
  const double r = PI.r;
  const double dA = r * G.w * PI.deltaPerp;
  double primL[NQ], primR[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    primL[q] = m.prim[q * m.Ntot + cL] + m.grad[q * m.Ntot + cL] * PI.deltaL + m.gradp[q * m.Ntot + cL] * dpL;
    primR[q] = m.prim[q * m.Ntot + cR] - m.grad[q * m.Ntot + cR] * PI.deltaR - m.gradp[q * m.Ntot + cR] * dpR;
  }
  ViscIn vi;
  if ((FEAT & F_VISC) && P.visc > 0.0) {
    visc_cells(m, vi, cL, cR);
    vi.cm = cm;
    const double4 rot = ann_rotation<FEAT>(m, P, PI.ann_in);
    if (PI.dir == 0) {
      vi.r_cell_L = r - PI.deltaL;
      vi.r_cell_R = r + PI.deltaR;
      vi.rdr_om = rot.w;  // at the face between the annuli
    } else {
      vi.r_cell_L = r;
      vi.r_cell_R = r;
      vi.rdr_om = rot.y;  // z faces sit at the annulus centre
    }
  }
  Y[0] = Y[1] = Y[2] = 0.0;
  return d_face_flux<NQ, 0, FEAT>(primL, primR, PI.dir == 0 ? 0 : 2, r, dA, 0.0, dt, P, sc, &vi, X, Y);
}

// every r (or z) overlap face once: its increments to Fface
template <int NQ, int FEAT>
__global__ __launch_bounds__(TILE, FLUX_WAVES) void k_flux_rz(DevMesh m, const int2 *ftiles, double fac) {
  const int2 tl = ftiles[tile_of(m)];
  const int p = tl.x;
  const int f = tl.y + threadIdx.x;
  const PairState S = m.pst[p];
  if (f >= S.count) return;
  const PairInfo PI = m.pairs[p];
  const DiskParams &P = disk_of(m, PI.disk);
  const DiskScalars &sc = m.sc[PI.disk];
  const long long g = PI.face_base + f;
  double X[NQ], Y[3];
  int e = rz_face_increments<NQ, FEAT>(m, PI, P, sc, fac * sc.dt, face_geo(m, PI, S, g), X, Y);
  if (e) set_err(m.err, e);
#pragma unroll
  for (int q = 0; q < NQ; ++q) m.Fface[q * m.NFtot + g] = X[q];
  if ((FEAT & F_VISC) && P.visc > 0.0) {
#pragma unroll
    for (int v = 0; v < 3; ++v) m.Fface[(NQ + v) * m.NFtot + g] = Y[v];
  }
}

// ----------------------------------------------------------------------------
// Per-cell update of one substep:
//   cell_copy (first substep), cell_adjust_RK_cons, the flux increments in
//   the reference's order, cell_add_src, cell_add_visc_src,
//   cell_add_split_fictitious, cell_update_phi, cell_update_dphi and
//   cell_calc_prim (src/TimeStep/timestep_routines.c:31-82).  Conserved
//   variables at the start of the substep are prim2cons of the primitives,
//   which is what cell_calc_cons left there (:112).
// ----------------------------------------------------------------------------
// the increments of a face evaluated in place (fused schedule): the same
// additions as add_face_R / add_face_L from the stored ones
template <int NQ>
__device__ __forceinline__ void add_inc_R(const double *X, const double *Y, double *u, double r_cell, bool visc) {
#pragma unroll
  for (int q = 0; q < NQ; ++q) u[q] += X[q];
  if (visc) {
    u[SRR] += Y[0] / r_cell;
    u[LLL] += Y[1];
    u[TAU] += Y[2];
  }
}
template <int NQ>
__device__ __forceinline__ void add_inc_L(const double *X, const double *Y, double *u, double r_cell, bool visc) {
#pragma unroll
  for (int q = 0; q < NQ; ++q) u[q] += -X[q];
  if (visc) {
    u[SRR] += -Y[0] / r_cell;
    u[LLL] += -Y[1];
    u[TAU] += -Y[2];
  }
}
template <int NQ>
__device__ __forceinline__ void add_face_R(const double *F, long long stride, long long g,
                                           double *u, double r_cell, bool visc) {
#pragma unroll
  for (int q = 0; q < NQ; ++q) u[q] += F[q * stride + g];
  if (visc) {
    u[SRR] += F[NQ * stride + g] / r_cell;
    u[LLL] += F[(NQ + 1) * stride + g];
    u[TAU] += F[(NQ + 2) * stride + g];
  }
}
template <int NQ>
__device__ __forceinline__ void add_face_L(const double *F, long long stride, long long g,
                                           double *u, double r_cell, bool visc) {
#pragma unroll
  for (int q = 0; q < NQ; ++q) u[q] += -F[q * stride + g];
  if (visc) {
    u[SRR] += -F[NQ * stride + g] / r_cell;
    u[LLL] += -F[(NQ + 1) * stride + g];
    u[TAU] += -F[(NQ + 2) * stride + g];
  }
}

// new tiph of one cell, cell_update_phi, src/Cell/cell_misc.c:59-79
__device__ __forceinline__ double new_tiph(double traw, double rkt_in, bool copy_first, double RK, double wtot,
                                           double dt, double r, double *rkt_out) {
  double tb = clean_pi(traw);
  const double rk0 = copy_first ? traw : rkt_in;
  // the unwrap loops of cell_misc.c:65-66, one trip by selects and the loops
  // from the start only when that trip does not end them (as wrap_range)
  double rkt = tb - rk0 > HALF_TURN ? rk0 + TWO_PI : (tb - rk0 < -HALF_TURN ? rk0 - TWO_PI : rk0);
  if (tb - rkt > HALF_TURN || tb - rkt < -HALF_TURN) {
    rkt = rk0;
    int n = 0;
#pragma unroll 1
    while (tb - rkt > HALF_TURN && n++ < MAXWRAP) rkt += TWO_PI;
#pragma unroll 1
    while (tb - rkt < -HALF_TURN && n++ < MAXWRAP) rkt -= TWO_PI;
  }
  *rkt_out = rkt;
  double t = (1.0 - RK) * tb + RK * rkt;
  t += wtot * dt / r;
  return t;
}


// add the stored increments of faces [b, e) of one range in face order, as
// add_face_R (R) or add_face_L; two faces' loads are issued before their adds
template <int NQ, bool R>
__device__ __forceinline__ void add_faces(const double *F, long long stride, long long base, int b, int e, double *u,
                                          double r_cell, bool visc) {
  constexpr int NV = NQ + 3;
  int f = b;
  for (; f + 1 < e; f += 2) {
    double x0[NV], x1[NV];
    const long long g = base + f;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      x0[q] = F[q * stride + g];
      x1[q] = F[q * stride + g + 1];
    }
    if (visc) {
#pragma unroll
      for (int v = NQ; v < NV; ++v) {
        x0[v] = F[v * stride + g];
        x1[v] = F[v * stride + g + 1];
      }
    }
    if (R) {
      add_inc_R<NQ>(x0, x0 + NQ, u, r_cell, visc);
      add_inc_R<NQ>(x1, x1 + NQ, u, r_cell, visc);
    } else {
      add_inc_L<NQ>(x0, x0 + NQ, u, r_cell, visc);
      add_inc_L<NQ>(x1, x1 + NQ, u, r_cell, visc);
    }
  }
  if (f < e) {
    if (R)
      add_face_R<NQ>(F, stride, base + f, u, r_cell, visc);
    else
      add_face_L<NQ>(F, stride, base + f, u, r_cell, visc);
  }
}

// FUSE: FUSE_R / FUSE_Z evaluate the r / z faces' increments in place (each
// face by both of its cells) instead of gathering the stored ones
enum { FUSE_R = 1, FUSE_Z = 2 };
template <int FEAT>
constexpr int update_waves() {
  return FEAT == F_HLL ? UPDATE_WAVES_HLL : UPDATE_WAVES;
}
template <int NQ, int FEAT, int FUSE>
__global__ __launch_bounds__(TILE, update_waves<FEAT>()) void k_update(DevMesh m, const int2 *tiles, double fac,
                                                                       int copy_first) {
  const int ti = tile_of(m);
  const int2 tl = tiles[ti];
  const AnnInfo A = m.ann[tl.x];
  const int t = threadIdx.x;
  const int n = min(TILE - 1, A.N - tl.y);  // cells of this tile
  const int j = tl.y + t;
  __shared__ double s_md[2][TILE / 64];
  __shared__ double s_F[NQ + 3][TILE];      // increments of phi faces j0-1 .. j0+n-1
  double md0 = 0.0, md1 = 0.0, tq = 0.0;
  const DiskParams &P = disk_of(m, A.disk);
  const DiskScalars &sc = m.sc[A.disk];
  const double4 rot = ann_rotation<FEAT>(m, P, tl.x);
  // the cell's r- and z-face ranges (the heads of the update's dependent
  // gathers), loaded before the phi faces so that their latency hides behind
  // the Riemann solves: as R, up to two runs of pair (i-1, i) / (k-1, k); as
  // L, one range of pair (i, i+1) / (k, k+1)
  int rb[3], re[3], zb[3] = {0, 0, 0}, ze[3] = {0, 0, 0};
  {
    const long long cz = A.zbase + j;
    const bool live = t < n;
    cell_face_ranges(m, live && A.i >= 1 ? rpair_of(P, A.k, A.i - 1) : -1,
                     live && A.i <= P.NR - 2 ? rpair_of(P, A.k, A.i) : -1, m.fo_r, cz, j, A.N, rb, re);
    if (P.NglobalZ != 1)
      cell_face_ranges(m, live && A.k >= 1 ? zpair_of(P, A.k - 1, A.i) : -1,
                       live && A.k <= P.NZ - 2 ? zpair_of(P, A.k, A.i) : -1, m.fo_z, cz, j, A.N, zb, ze);
  }
  // phi faces: each is solved once, by the thread of its left cell (thread 0
  // also solves the face shared with the previous tile)
  if (t <= n) {
    int jf = tl.y - 1 + t;
    if (jf < 0) jf += A.N;
    double X[NQ], Y[3];
    int e = phi_face_flux<NQ, FEAT>(m, A, tl.x, jf, P, sc, fac * sc.dt, rot.y, X, Y);
    if (e) set_err(m.err, e);
#pragma unroll
    for (int q = 0; q < NQ; ++q) s_F[q][t] = X[q];
#pragma unroll
    for (int v = 0; v < 3; ++v) s_F[NQ + v][t] = Y[v];
  }
  __syncthreads();
  if (t < n) {
    const double dt = fac * sc.dt;
    const double RK = sc.RK;
    const bool visc = (FEAT & F_VISC) && P.visc > 0.0;
    const long long c = A.zbase + j;
    const int N = A.N, k = A.k, i = A.i;
    const double rp = A.rp, rm = A.rm, zp = A.zp, zm = A.zm;
    const double r = .5 * (rp + rm);
    const double dz = zp - zm;
    const double dphi = m.dphi[c];
    double pr[NQ], u[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) pr[q] = m.prim[q * m.Ntot + c];
    // conserved state left by cell_calc_cons
    d_prim2cons<NQ>(pr, u, r, .5 * (rp * rp - rm * rm) * dphi * dz, P.gamma, P.NC);
    if (copy_first) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) m.RKcons[q * m.Ntot + c] = u[q];
    }
    // cell_adjust_RK_cons, cell_misc.c:43-57
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double rk = copy_first ? u[q] : m.RKcons[q * m.Ntot + c];
      u[q] = (1.0 - RK) * u[q] + RK * rk;
    }
    // phi faces (timestep_routines.c:42-51): face j-1 has this cell as R,
    // face j as L; for j = 0 the L face comes first.
    {
      auto phiR = [&]() {  // face (j-1, j): slot t
#pragma unroll
        for (int q = 0; q < NQ; ++q) u[q] += s_F[q][t];
        if (visc) {
          u[SRR] += s_F[NQ][t] / r;
          u[LLL] += s_F[NQ + 1][t];
          u[TAU] += s_F[NQ + 2][t];
        }
      };
      auto phiL = [&]() {  // face (j, j+1): slot t+1
#pragma unroll
        for (int q = 0; q < NQ; ++q) u[q] += -s_F[q][t + 1];
        if (visc) {
          u[SRR] += -s_F[NQ][t + 1] / r;
          u[LLL] += -s_F[NQ + 1][t + 1];
          u[TAU] += -s_F[NQ + 2][t + 1];
        }
      };
      if (j == 0) {
        phiL();
        phiR();
      } else {
        phiR();
        phiL();
      }
    }
    // r faces in face order: as R of pair (i-1, i), then as L of (i, i+1);
    // then the z faces the same way
#pragma unroll
    for (int dir = 0; dir < 2; ++dir) {
      if (dir == 1 && P.NglobalZ == 1) break;
      const bool fuse = dir == 0 ? (FUSE & FUSE_R) != 0 : (FUSE & FUSE_Z) != 0;
      const int *b = dir == 0 ? rb : zb, *e = dir == 0 ? re : ze;
      const bool lo = dir == 0 ? i >= 1 : k >= 1, hi = dir == 0 ? i <= P.NR - 2 : k <= P.NZ - 2;
      if (lo) {
        const int pq = dir == 0 ? rpair_of(P, k, i - 1) : zpair_of(P, k - 1, i);
        const PairInfo PI = m.pairs[pq];
        const double rc = dir == 0 ? PI.r + PI.deltaR : PI.r;
#pragma unroll
        for (int w = 0; w < 2; ++w) {
          if (fuse) {
            const PairState S = m.pst[pq];
            for (int f = b[w]; f < e[w]; ++f) {
              double X[NQ], Y[3];
              const int er =
                  rz_face_increments<NQ, FEAT>(m, PI, P, sc, dt, face_geo(m, PI, S, PI.face_base + f), X, Y);
              if (er) set_err(m.err, er);
              add_inc_R<NQ>(X, Y, u, rc, visc);
            }
          } else {
            add_faces<NQ, true>(m.Fface, m.NFtot, PI.face_base, b[w], e[w], u, rc, visc);
          }
        }
      }
      if (hi) {
        const int pq = dir == 0 ? rpair_of(P, k, i) : zpair_of(P, k, i);
        const PairInfo PI = m.pairs[pq];
        const double rc = dir == 0 ? PI.r - PI.deltaL : PI.r;
        if (fuse) {
          const PairState S = m.pst[pq];
          for (int f = b[2]; f < e[2]; ++f) {
            double X[NQ], Y[3];
            const int er =
                rz_face_increments<NQ, FEAT>(m, PI, P, sc, dt, face_geo(m, PI, S, PI.face_base + f), X, Y);
            if (er) set_err(m.err, er);
            add_inc_L<NQ>(X, Y, u, rc, visc);
          }
        } else {
          add_faces<NQ, false>(m.Fface, m.NFtot, PI.face_base, b[2], e[2], u, rc, visc);
        }
      }
    }
    // cell_add_src, cell_add_visc_src, cell_add_split_fictitious
    // (src/Cell/cell_source.c:73-277, cell_add_split_fictitious.c:11-44)
    const double traw = m.tiph[c];
    const double tcl = clean_pi(traw);
    {
      double phi = tcl - .5 * dphi;
      double cphi, sphi;
      sincos(phi, &sphi, &cphi);
      double dV = dphi * .5 * (rp * rp - rm * rm) * dz;
      double zc = .5 * (zp + zm);
      const double Fp =
          src_grav_sink_cool<FEAT>(P, sc, rot, i, k, r, dV, zc, cphi, sphi, pr, u, dt, &md0, &md1, nullptr);
      if (m.torque_tile && interior_cell(P, i, k)) tq += dt * dV * pr[RHO] * Fp * r;
      if (visc)
        src_visc<FEAT>(P, sc, rot, r, dV, cphi, sphi, pr, m.grad[UPP * m.Ntot + c], m.gradp[URR * m.Ntot + c], u,
                       dt);
    }
    src_fict<FEAT>(rot, r, dt, u);
    // cell_update_phi + cell_update_dphi (the left neighbour's new edge is
    // recomputed here from its old state)
    const double wtot = m.wiph[tl.x] + rot.x;
    double rkt_own, rkt_nb;
    double tnew = new_tiph(traw, m.RKtiph[c], copy_first, RK, wtot, dt, r, &rkt_own);
    const long long cm1 = A.zbase + (j == 0 ? N - 1 : j - 1);
    double tnew_m1 = new_tiph(m.tiph[cm1], m.RKtiph[cm1], copy_first, RK, wtot, dt, r, &rkt_nb);
    double dphi_new = wrap_p(tnew - tnew_m1);
    m.RKtiph[c] = rkt_own;
    m.tiph_new[c] = tnew;
    m.dphi_new[c] = dphi_new;
    if (m.u_save) {  // the conserved state cons2prim sees (recorded for the adjoint)
#pragma unroll
      for (int q = 0; q < NQ; ++q) m.u_save[q * m.Ntot + c] = u[q];
    }
    // cell_calc_prim, cell_conversion.c:105-122
    double out[NQ];
    d_cons2prim<NQ>(u, out, r, .5 * (rp * rp - rm * rm) * dphi_new * dz, P);
#pragma unroll
    for (int q = 0; q < NQ; ++q) m.prim_new[q * m.Ntot + c] = out[q];
  }
  if (m.torque_tile) {  // the tile's gravitational torque on the gas (adjoint loss)
    for (int off = 32; off > 0; off >>= 1) tq += __shfl_down(tq, off);
    if ((threadIdx.x & 63) == 0) s_md[0][threadIdx.x >> 6] = tq;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t0 = 0.0;
      for (int w = 0; w < TILE / 64; ++w) t0 += s_md[0][w];
      m.torque_tile[A.utile + tl.y / (TILE - 1)] = t0;
    }
    __syncthreads();
  }
  if ((FEAT & F_SINK) && P.sink == 1) {
    // the tile's accretion of each mass, summed in a fixed order into the
    // tile's own slot; k_scalar folds a disk's slots in tile order, so Mdot
    // does not depend on the order in which workgroups run
    for (int off = 32; off > 0; off >>= 1) {
      md0 += __shfl_down(md0, off);
      md1 += __shfl_down(md1, off);
    }
    if ((threadIdx.x & 63) == 0) {
      s_md[0][threadIdx.x >> 6] = md0;
      s_md[1][threadIdx.x >> 6] = md1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double t0 = 0.0, t1 = 0.0;
      for (int w = 0; w < TILE / 64; ++w) {
        t0 += s_md[0][w];
        t1 += s_md[1][w];
      }
      const int slot = A.utile + tl.y / (TILE - 1);
      m.mdot_tile[2 * slot] = t0;
      m.mdot_tile[2 * slot + 1] = t1;
    }
  }
}

// ----------------------------------------------------------------------------
// Outflow r boundary, cell_boundary_outflow_r, src/Cell/cell_boundary.c:14-139:
// the ghost annulus takes the overlap-area-weighted primitives of its
// neighbour.  stage 0: inner ghost (i = 0) and outer ghost NR-2; stage 1:
// outer ghost NR-1 (from the new NR-2).
// ----------------------------------------------------------------------------
// the new primitives of cell c2 of outer ghost annulus i2 = NR-2 (stage 0
// below: the overlap-area-weighted remap of annulus NR-3), with stage 0's
// arithmetic, for stage 1 when both run in one launch
template <int NQ>
__device__ __forceinline__ void outer_ghost_value(const DevMesh &m, const DiskParams &P, int k, int i2, int c2,
                                                  double *w) {
  const AnnInfo A2 = m.ann[P.ann_base + k * P.NR + i2];
#pragma unroll
  for (int q = 0; q < NQ; ++q) w[q] = 0.0;
  faces_as_R(m, rpair_of(P, k, i2 - 1), (int)(c2 - A2.zbase), A2.N, [&](const PairInfo &PI, long long g) {
    const int cL = m.fL[g];
    double dA = PI.r * m.fdphi[g] * PI.deltaPerp;
#pragma unroll
    for (int q = 0; q < NQ; ++q) w[q] += m.prim[q * m.Ntot + cL] * dA;
  });
  const double dA = (A2.zp - A2.zm) * A2.rm * m.dphi[c2];
#pragma unroll
  for (int q = 0; q < NQ; ++q) w[q] = w[q] / dA;
}

// nested: stage 1's annulus NR-1 runs in the same launch as stage 0 and
// recomputes the NR-2 values it reads instead of loading them
template <int NQ>
__device__ __forceinline__ void bc_r_tile(const DevMesh &m, const int2 tl, bool nested) {
  const AnnInfo A = m.ann[tl.x];
  const int j = tl.y + threadIdx.x;
  if (j >= A.N) return;
  const DiskParams &P = disk_of(m, A.disk);
  const long long c = A.zbase + j;
  const int k = A.k, i = A.i;
  const double dz = A.zp - A.zm;
  double v[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) v[q] = 0.0;
  double r_face;
  if (i == 0) {
    // cL->prim += (cR->prim - 0*grad*(...)/2.)*dA over the pair's faces
    const double span = m.rf[P.geo_r + i + 2] - m.rf[P.geo_r + i];
    r_face = A.rp;
    faces_as_L(m, rpair_of(P, k, 0), m.fo_r, c, j, A.N, [&](const PairInfo &PI, long long g) {
      const int cR = m.fR[g];
      double dA = PI.r * m.fdphi[g] * PI.deltaPerp;
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        v[q] += (m.prim[q * m.Ntot + cR] - 0 * m.grad[q * m.Ntot + cR] * span / 2.) * dA;
    });
  } else {
    r_face = A.rm;
    const bool from_ghost = nested && i == P.NR - 1;
    faces_as_R(m, rpair_of(P, k, i - 1), j, A.N, [&](const PairInfo &PI, long long g) {
      const int cL = m.fL[g];
      double dA = PI.r * m.fdphi[g] * PI.deltaPerp;
      if (from_ghost) {
        double w[NQ];
        outer_ghost_value<NQ>(m, P, k, i - 1, cL, w);
#pragma unroll
        for (int q = 0; q < NQ; ++q) v[q] += w[q] * dA;
      } else {
#pragma unroll
        for (int q = 0; q < NQ; ++q) v[q] += m.prim[q * m.Ntot + cL] * dA;
      }
    });
  }
  const double dA = dz * r_face * m.dphi[c];
#pragma unroll
  for (int q = 0; q < NQ; ++q) m.prim[q * m.Ntot + c] = v[q] / dA;
}
template <int NQ>
__global__ __launch_bounds__(TILE) void k_bc_r(DevMesh m, const int2 *tiles) {
  bc_r_tile<NQ>(m, tiles[tile_of(m)], false);
}

// z halo on one rank: periodic copy of (prim, RKcons, tiph) into the ghost
// layers, src/Cell/cell_sync.c:156-205.
template <int NQ>
__global__ __launch_bounds__(TILE) void k_sync_z(DevMesh m, const int2 *tiles) {
  const int2 tl = tiles[tile_of(m)];
  const AnnInfo A = m.ann[tl.x];
  const int j = tl.y + threadIdx.x;
  if (j >= A.N) return;
  const DiskParams &P = disk_of(m, A.disk);
  const int k = A.k;
  int ksrc;
  if (k < P.NgMinZ)
    ksrc = P.NZ - 2 * P.NgMaxZ + k;
  else
    ksrc = P.NgMinZ + (k - (P.NZ - P.NgMaxZ));
  const long long c = A.zbase + j;
  const long long s = c + (long long)(ksrc - k) * P.NL;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    m.prim[q * m.Ntot + c] = m.prim[q * m.Ntot + s];
    m.RKcons[q * m.Ntot + c] = m.RKcons[q * m.Ntot + s];
  }
  m.tiph_new[c] = m.tiph_new[s];
}

// ----------------------------------------------------------------------------
// Fixed r boundaries, cell_boundary_fixed_r (src/Cell/cell_boundary.c:208-232),
// then boundary damping, cell_bc_damp (src/Cell/cell_misc.c:98-132), on the
// annuli the host marked in their InitRow.  A ghost cell of a fixed boundary
// takes the single-cell initial data (src/Cell/initial_data/*.c) at its new
// edge; a damped cell relaxes towards the initial data by 1 - exp(-rate dt).
// Per cell the boundary comes first, as in the reference.
// ----------------------------------------------------------------------------
template <int NQ>
__global__ __launch_bounds__(TILE) void k_bc_init(DevMesh m, const int2 *tiles, double fac) {
  const int2 tl = tiles[tile_of(m)];
  const AnnInfo A = m.ann[tl.x];
  const int j = tl.y + threadIdx.x;
  if (j >= A.N) return;
  const InitRow R = m.init[tl.x];
  const DiskParams &P = disk_of(m, A.disk);
  const long long c = A.zbase + j;
  const double r = 0.5 * (A.rm + A.rp);
  double pr[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) pr[q] = m.prim[q * m.Ntot + c];
  if (R.flags & INIT_FIXED) single_init<NQ>(P.initial_data_type, R, r, m.tiph[c], m.dphi[c], pr);
  if (R.flags & INIT_DAMP) {  // NUM_Q = 5 and initial data of r alone (the host checks)
    const double dt = fac * m.sc[A.disk].dt;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const double dprim = R.prim[q] - pr[q];
      pr[q] += (1. - exp(-R.rate * dt)) * dprim;
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) m.prim[q * m.Ntot + c] = pr[q];
}

// SC_DT_FOLD: dtmin of disk d = min(dtmin, the positive tile minima below
// 1e100) over the disk's interior tiles; every thread of the workgroup calls it.
template <int NT>
__device__ __forceinline__ void fold_dt(const DevMesh &m, const DiskParams &P, DiskScalars &s) {
  __shared__ double s_m[NT / 64];
  double v = 1.e100;
  // four independent loads in flight per thread
  for (int it = P.itile0 + (int)threadIdx.x; it < P.itile1; it += 4 * NT) {
    double b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int iu = it + u * NT;
      b[u] = iu < P.itile1 ? m.dt_tile[iu] : 1.e100;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (b[u] > 0.0 && b[u] < v) v = b[u];
  }
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_down(v, off);
    if (o < v) v = o;
  }
  if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = s_m[0];
    for (int w = 1; w < NT / 64; ++w)
      if (s_m[w] < b) b = s_m[w];
    if (b < 1.e100) {
      const unsigned long long bits = (unsigned long long)__double_as_longlong(b);
      if (bits < s.dtmin_bits) s.dtmin_bits = bits;
    }
  }
}

// ----------------------------------------------------------------------------
// Per-disk scalar updates (time, dt, the two point masses):
// src/TimeStep/timestep_algorithm.c:11-41, src/TimeStep/timestep_set.c:16-21,
// src/GravMass/GravMass_misc.c:9-77.
// ----------------------------------------------------------------------------
__device__ __forceinline__ void d_gm_move(const DiskParams &P, DiskScalars &s, double t, double dt) {
  double m0 = s.gm[0].M, m1 = s.gm[1].M;
  double M = m0 + m1;
  double a_0 = 1.0, tau_0 = P.orb_tscale, T0 = P.orb_t0;
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
  s.gm[0].phi += omega * dt;
  s.gm[1].phi += omega * dt;
  s.gm[0].r = r0;
  s.gm[1].r = r1;
  s.gm[0].omega = omega;
  s.gm[1].omega = omega;
}

// launch-uniform gravity terms of the current mass positions (computed
// with the same device functions the per-cell code would call)
__device__ __forceinline__ void d_mass_terms(const DiskParams &P, DiskScalars &s) {
  for (int p = 0; p < 2; ++p) {
    s.xbh[p] = s.gm[p].r * cos(s.gm[p].phi);
    s.ybh[p] = s.gm[p].r * sin(s.gm[p].phi);
  }
  s.epsn = pow(P.g_eps, P.phi_order);
}

__device__ __forceinline__ void scalar_ops(const DiskParams &P, DiskScalars &s, int ops) {
  if (ops & SC_STEP_BEGIN) {  // timestep_set_dt, gravMass_copy, RK = 0
    double dt = __longlong_as_double((long long)s.dtmin_bits);
    if (s.t + dt > P.t_max) dt = P.t_max - s.t;
    s.dt = dt;
    for (int p = 0; p < 2; ++p) {
      s.gm[p].RK_r = s.gm[p].r;
      s.gm[p].RK_phi = s.gm[p].phi;
      s.gm[p].RK_M = s.gm[p].M;
      s.gm[p].RK_omega = s.gm[p].omega;
    }
    s.RK = 0.0;
  }
  if (ops & SC_AFTER_UPDATE) {  // Mdot of cell_add_src, then gravMass_update_RK
    s.gm[0].Mdot = s.mdot_acc[0];
    s.gm[1].Mdot = s.mdot_acc[1];
    double RK = s.RK;
    for (int p = 0; p < 2; ++p) {
      GravState &g = s.gm[p];
      int n = 0;
      while (g.phi - g.RK_phi > M_PI && n++ < MAXWRAP) g.RK_phi += 2. * M_PI;
      while (g.phi - g.RK_phi < -M_PI && n++ < MAXWRAP) g.RK_phi -= 2. * M_PI;
      g.r = (1.0 - RK) * g.r + RK * g.RK_r;
      g.phi = (1.0 - RK) * g.phi + RK * g.RK_phi;
      g.M = (1.0 - RK) * g.M + RK * g.RK_M;
      g.omega = (1.0 - RK) * g.omega + RK * g.RK_omega;
    }
  }
  if (ops & SC_SUB1_END) {  // gravMass_move(t, 1.0*dt); RK = 0.5
    d_gm_move(P, s, s.t, 1.0 * s.dt);
    s.RK = 0.5;
  }
  if (ops & SC_SUB2_END) {  // gravMass_move(t, 0.5*dt); t += dt; Macc += Mdot*dt
    d_gm_move(P, s, s.t, 0.5 * s.dt);
    s.t += s.dt;
    for (int p = 0; p < 2; ++p) s.gm[p].Macc = s.gm[p].Macc + s.gm[p].Mdot * s.dt;
    s.dtmin_bits = (unsigned long long)__double_as_longlong(1.e100);
    d_mass_terms(P, s);
  }
  if (ops & SC_EULER_END) {  // timestep_forward_euler: gravMass_move(t, dt); t += dt
    d_gm_move(P, s, s.t, 1.0 * s.dt);
    s.t += s.dt;
    s.dtmin_bits = (unsigned long long)__double_as_longlong(1.e100);
    d_mass_terms(P, s);
  }
  if (ops & SC_SUB_BEGIN) {  // gravMass_clean_pi
    for (int p = 0; p < 2; ++p) {
      double phi = s.gm[p].phi;
      int n = 0;
      while (phi > 2. * M_PI && n++ < MAXWRAP) phi -= 2. * M_PI;
      while (phi < 0.0 && n++ < MAXWRAP) phi += 2. * M_PI;
      s.gm[p].phi = phi;
    }
    s.mdot_acc[0] = 0.0;
    s.mdot_acc[1] = 0.0;
    d_mass_terms(P, s);
  }
  if (ops & SC_RESET_DT) s.dtmin_bits = (unsigned long long)__double_as_longlong(1.e100);
  if (ops & SC_MASS) d_mass_terms(P, s);
}

// SC_AFTER_UPDATE with a sink: mdot_acc of disk d = the sum of its update
// tiles' accretion slots (k_update), in a fixed order: thread t adds tiles
// t, t + NT, ... in turn, then the waves' sums are added in order.
template <int NT>
__device__ __forceinline__ void fold_mdot(const DevMesh &m, const DiskParams &P, DiskScalars &s) {
  __shared__ double s_a[2][NT / 64];
  double v0 = 0.0, v1 = 0.0;
  for (int it = P.utile0 + (int)threadIdx.x; it < P.utile1; it += NT) {
    v0 += m.mdot_tile[2 * it];
    v1 += m.mdot_tile[2 * it + 1];
  }
  for (int off = 32; off > 0; off >>= 1) {
    v0 += __shfl_down(v0, off);
    v1 += __shfl_down(v1, off);
  }
  if ((threadIdx.x & 63) == 0) {
    s_a[0][threadIdx.x >> 6] = v0;
    s_a[1][threadIdx.x >> 6] = v1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = 0.0, t1 = 0.0;
    for (int w = 0; w < NT / 64; ++w) {
      t0 += s_a[0][w];
      t1 += s_a[1][w];
    }
    s.mdot_acc[0] = t0;
    s.mdot_acc[1] = t1;
  }
  __syncthreads();
}

// SC_AFTER_UPDATE with torque_acc: torque_acc[d] += the sum of the disk's
// update tiles' torque (k_update), thread t adding tiles t, t + NT, ... in
// turn, then the waves' sums in order
template <int NT>
__device__ __forceinline__ void fold_torque(const DevMesh &m, const DiskParams &P, int d) {
  __shared__ double s_a[NT / 64];
  double v = 0.0;
  for (int it = P.utile0 + (int)threadIdx.x; it < P.utile1; it += NT) v += m.torque_tile[it];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  if ((threadIdx.x & 63) == 0) s_a[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < NT / 64; ++w) t += s_a[w];
    m.torque_acc[d] += t;
  }
}

// the scalar operations of disk d, by one workgroup of NT threads
template <int NT>
__device__ __forceinline__ void scalar_block(const DevMesh &m, int d, int ops) {
  const DiskParams &P = m.dp[d];
  DiskScalars &s = m.sc[d];
  if (m.sc_rec && (ops & SC_STEP_BEGIN) && threadIdx.x == 0) m.sc_rec[3 * d] = s;  // (before the folds write s)
  if (ops & (SC_DT_FOLD | SC_STEP_BEGIN)) fold_dt<NT>(m, P, s);
  if ((ops & SC_AFTER_UPDATE) && P.sink == 1) fold_mdot<NT>(m, P, s);
  if ((ops & SC_AFTER_UPDATE) && m.torque_acc) fold_torque<NT>(m, P, d);
  if (threadIdx.x == 0) {
    // on register copies: the operations' dependent chains would otherwise
    // go through global memory (the two structs may alias)
    const DiskParams Pl = P;
    DiskScalars sl = s;
    scalar_ops(Pl, sl, ops);
    s = sl;
    if (m.sc_rec && (ops & SC_STEP_BEGIN)) m.sc_rec[3 * d + 1] = sl;
    if (m.sc_rec && (ops & SC_SUB1_END)) m.sc_rec[3 * d + 2] = sl;
  }
  // the A_MILOS frame rotation (src/Sim/sim_w_a.c) of every annulus, at the
  // mass separation a the next kernels will see: after any mass move (the
  // other frames do not depend on a: the host tabulates them once)
  if ((ops & (SC_SUB2_END | SC_EULER_END | SC_SUB_BEGIN | SC_MASS)) && P.w_analytic == A_MILOS) {
    __syncthreads();
    const double a = s.gm[0].r + s.gm[1].r;
    const int na = P.NR * P.NZ;
    for (int k = threadIdx.x; k < na; k += NT) {
      const AnnInfo &A = m.ann[P.ann_base + k];
      const double r = .5 * (A.rp + A.rm);
      m.ann_rot[P.ann_base + k] =
          make_double4(d_rOm_a(P.w_analytic, r, a), d_rdrOm_a(P.w_analytic, r, a),
                       d_dtOm_a(P.w_analytic, r, a, P.orb_tscale), d_rdrOm_a(P.w_analytic, A.rp, a));
    }
  }
}
__global__ __launch_bounds__(SCALAR_THREADS) void k_scalar(DevMesh m, int ndisk, int ops, double arg) {
  scalar_block<SCALAR_THREADS>(m, blockIdx.x, ops);  // one workgroup per disk
}

// The outflow boundary (both stages: annulus NR-1 recomputes the NR-2
// values it needs) and, in the workgroups past its tiles, the post-update
// scalar operations of every disk, in one launch: the parts are independent
// (the boundary remap reads no scalar, the scalar operations touch no cell)
// and each alone is a few microseconds of launch latency.  ops must not fold
// the CFL minimum (a fold wants SCALAR_THREADS).
template <int NQ>
__global__ __launch_bounds__(TILE) void k_bc_r_scalar(DevMesh m, const int2 *tiles, int ntiles, int ops) {
  const int b = tile_of(m);
  if (b < ntiles)
    bc_r_tile<NQ>(m, tiles[b], true);
  else
    scalar_block<TILE>(m, b - ntiles, ops);
}

// ----------------------------------------------------------------------------
// host-side launchers
// ----------------------------------------------------------------------------
KProf *g_kprof = nullptr;

// overlap faces of the r (dir 0) or z (dir 1) pairs
// skip: regular pairs' faces need not be stored (the step's own kernels
// compute them)
template <int NQ>
static void launch_faces_q(const DevMesh &m, const LaunchPlan &L, hipStream_t st, int dir, int skip) {
  const int np = dir == 0 ? L.n_rpairs : L.n_zpairs;
  const int p0 = dir == 0 ? 0 : L.n_rpairs;
  if (!np) return;
  const int nst = dir == 0 ? L.stage_r : L.stage_z;
  const size_t lds = (size_t)nst * 2 * sizeof(double);
  const char *nm = dir == 0 ? "k_faces_r" : "k_faces_z";
  if (L.face_threads == 256)
    KLAUNCH(st, nm, hipLaunchKernelGGL(k_faces<256>, dim3(np), dim3(256), lds, st, m, p0, nst, skip));
  else if (L.face_threads == 512)
    KLAUNCH(st, nm, hipLaunchKernelGGL(k_faces<512>, dim3(np), dim3(512), lds, st, m, p0, nst, skip));
  else
    KLAUNCH(st, nm, hipLaunchKernelGGL(k_faces<1024>, dim3(np), dim3(1024), lds, st, m, p0, nst, skip));
}

template <int NQ, int FEAT>
static void launch_flux_q(const DevMesh &m, const int2 *ft, int nft, hipStream_t st, double fac, const char *name) {
  KLAUNCH(st, name, hipLaunchKernelGGL((k_flux_rz<NQ, FEAT>), dim3(nft), dim3(TILE), 0, st, m, ft, fac));
}

template <int NQ, int FEAT>
static void launch_update_q(const DevMesh &m, const LaunchPlan &L, hipStream_t st, double fac, int copy_first,
                            int fuse) {
  const dim3 G(L.n_utiles), B(TILE);
  if (fuse == (FUSE_R | FUSE_Z))
    KLAUNCH(st, "k_update", hipLaunchKernelGGL((k_update<NQ, FEAT, FUSE_R | FUSE_Z>), G, B, 0, st, m, L.utiles, fac, copy_first));
  else if (fuse == FUSE_R)
    KLAUNCH(st, "k_update", hipLaunchKernelGGL((k_update<NQ, FEAT, FUSE_R>), G, B, 0, st, m, L.utiles, fac, copy_first));
  else if (fuse == FUSE_Z)
    KLAUNCH(st, "k_update", hipLaunchKernelGGL((k_update<NQ, FEAT, FUSE_Z>), G, B, 0, st, m, L.utiles, fac, copy_first));
  else
    KLAUNCH(st, "k_update", hipLaunchKernelGGL((k_update<NQ, FEAT, 0>), G, B, 0, st, m, L.utiles, fac, copy_first));
}

template <int NQ, int FEAT>
static void launch_substep_q(const DevMesh &m, const LaunchPlan &L, hipStream_t st, double fac, int copy_first,
                             int recompute_w, int pre_ops, int post_ops) {
  const dim3 B(TILE);
  const dim3 SD(L.ndisk), S64(SCALAR_THREADS);
  if (pre_ops & SC_PLM_DT) {  // the CFL tile minima ride on the phi slopes; the fold follows them
    KLAUNCH(st, "k_plm_p_dt",
            hipLaunchKernelGGL((k_plm_p<NQ, true, FEAT>), dim3(L.n_ztiles), B, 0, st, m, L.ztiles));
    KLAUNCH(st, "k_scalar", hipLaunchKernelGGL(k_scalar, SD, S64, 0, st, m, L.ndisk, pre_ops & ~SC_PLM_DT, 0.0));
  } else {
    if (pre_ops) KLAUNCH(st, "k_scalar", hipLaunchKernelGGL(k_scalar, SD, S64, 0, st, m, L.ndisk, pre_ops, 0.0));
    KLAUNCH(st, "k_plm_p", hipLaunchKernelGGL((k_plm_p<NQ, false>), dim3(L.n_ztiles), B, 0, st, m, L.ztiles));
  }
  // the adjoint's recording forward (m.ftype set) keeps every face
  const int skip = m.ftype == nullptr;
  launch_faces_q<NQ>(m, L, st, 0, skip);
  launch_faces_q<NQ>(m, L, st, 1, skip);
  if (recompute_w) KLAUNCH(st, "k_set_w", hipLaunchKernelGGL(k_set_w, dim3(L.nann), dim3(64), 0, st, m, 0));
  // 3D: both slope directions in one pass, the r slopes kept apart for the r fluxes
  const bool both = L.n_zpairs && m.grad_r;
  if (both)
    KLAUNCH(st, "k_plm_rz", hipLaunchKernelGGL(k_plm_rz2<NQ>, dim3(L.n_ztiles), B, 0, st, m, L.ztiles, m.grad_r));
  else
    KLAUNCH(st, "k_plm_r", hipLaunchKernelGGL(k_plm_rz<NQ>, dim3(L.n_ztiles), B, 0, st, m, L.ztiles, 0));
  // the fused schedule evaluates the r faces inside k_update, which then
  // needs the r slopes: only without z faces (the z slopes overwrite them)
  const int fuse = (L.n_zpairs ? 0 : (L.fuse & FUSE_R)) | (L.fuse & FUSE_Z);
  if (L.n_rftiles && !(fuse & FUSE_R)) {
    DevMesh mr = m;
    if (both) mr.grad = m.grad_r;
    launch_flux_q<NQ, FEAT>(mr, L.rftiles, L.n_rftiles, st, fac, "k_flux_r");
  }
  if (L.n_zpairs) {
    if (!both) KLAUNCH(st, "k_plm_z", hipLaunchKernelGGL(k_plm_rz<NQ>, dim3(L.n_ztiles), B, 0, st, m, L.ztiles, 1));
    if (!(fuse & FUSE_Z)) launch_flux_q<NQ, FEAT>(m, L.zftiles, L.n_zftiles, st, fac, "k_flux_z");
  }
  launch_update_q<NQ, FEAT>(m, L, st, fac, copy_first, fuse);
  DevMesh m2 = m;  // after the update the new state lives in the *_new buffers
  m2.tiph = m.tiph_new;
  m2.prim = m.prim_new;
  m2.dphi = m.dphi_new;
  // the two boundary stages' tile lists are adjacent in the tile array
  const bool one_launch = L.n_bc0 && L.bc1 == L.bc0 + L.n_bc0 && !(post_ops & (SC_DT_FOLD | SC_STEP_BEGIN));
  if (one_launch) {
    const int nt = L.n_bc0 + L.n_bc1;
    KLAUNCH(st, "k_bc_r_scalar", hipLaunchKernelGGL(k_bc_r_scalar<NQ>, dim3(nt + (post_ops ? L.ndisk : 0)), B, 0, st,
                                                    m2, L.bc0, nt, post_ops));
  } else {
    if (post_ops) KLAUNCH(st, "k_scalar", hipLaunchKernelGGL(k_scalar, SD, S64, 0, st, m, L.ndisk, post_ops, 0.0));
    if (L.n_bc0) KLAUNCH(st, "k_bc_r", hipLaunchKernelGGL(k_bc_r<NQ>, dim3(L.n_bc0), B, 0, st, m2, L.bc0));
    if (L.n_bc1) KLAUNCH(st, "k_bc_r", hipLaunchKernelGGL(k_bc_r<NQ>, dim3(L.n_bc1), B, 0, st, m2, L.bc1));
  }
  if (L.n_binit)
    KLAUNCH(st, "k_bc_init", hipLaunchKernelGGL(k_bc_init<NQ>, dim3(L.n_binit), B, 0, st, m2, L.binit, fac));
  if (L.n_sz) KLAUNCH(st, "k_sync_z", hipLaunchKernelGGL(k_sync_z<NQ>, dim3(L.n_sz), B, 0, st, m2, L.sz));
}


// The compiled physics variants (feature sets, F_*): NUM_Q 5 has a variant
// per common configuration, every NUM_Q the general one.  The plan's feat is
// the union of what its disks need.
static const int FEAT_VARIANTS[] = {F_HLLC, F_HLL, F_VISC | F_HLLC, F_ALL};
int feature_variant(int need, int NQ) {
  if (NQ == 5)
    for (int v : FEAT_VARIANTS)
      if ((need & ~v) == 0) return v;
  return F_ALL;
}

template <int NQ>
static void launch_substep_nq(const DevMesh &m, const LaunchPlan &L, hipStream_t st, double fac, int copy_first,
                              int recompute_w, int pre_ops, int post_ops) {
  const int v = feature_variant(L.feat, NQ);
  if (NQ == 5 && v == F_HLLC)
    launch_substep_q<NQ, F_HLLC>(m, L, st, fac, copy_first, recompute_w, pre_ops, post_ops);
  else if (NQ == 5 && v == F_HLL)
    launch_substep_q<NQ, F_HLL>(m, L, st, fac, copy_first, recompute_w, pre_ops, post_ops);
  else if (NQ == 5 && v == (F_VISC | F_HLLC))
    launch_substep_q<NQ, F_VISC | F_HLLC>(m, L, st, fac, copy_first, recompute_w, pre_ops, post_ops);
  else
    launch_substep_q<NQ, F_ALL>(m, L, st, fac, copy_first, recompute_w, pre_ops, post_ops);
}

void launch_substep(const DevMesh &m, const LaunchPlan &L, hipStream_t st, int NQ, double fac, int copy_first,
                    int recompute_w, int pre_ops, int post_ops) {
  DISCO_NQ_DISPATCH(NQ, launch_substep_nq<Q>(m, L, st, fac, copy_first, recompute_w, pre_ops, post_ops));
}

bool prepare_launches(const LaunchPlan &L) {
  const int nst = L.stage_r > L.stage_z ? L.stage_r : L.stage_z;
  if (!nst) return true;
  const int bytes = (int)(nst * 2 * sizeof(double));
  return hipFuncSetAttribute((const void *)k_faces<1024>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) ==
             hipSuccess &&
         hipFuncSetAttribute((const void *)k_faces<512>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) ==
             hipSuccess &&
         hipFuncSetAttribute((const void *)k_faces<256>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) ==
             hipSuccess;
}

void launch_scalar(const DevMesh &m, const LaunchPlan &L, hipStream_t st, int op) {
  KLAUNCH(st, "k_scalar", hipLaunchKernelGGL(k_scalar, dim3(L.ndisk), dim3(SCALAR_THREADS), 0, st, m, L.ndisk, op, 0.0));
}

void launch_set_w(const DevMesh &m, const LaunchPlan &L, hipStream_t st) {
  KLAUNCH(st, "k_set_w", hipLaunchKernelGGL(k_set_w, dim3(L.nann), dim3(64), 0, st, m, 0));
}

void launch_mindt(const DevMesh &m, const LaunchPlan &L, hipStream_t st, int NQ) {
  if (!L.n_itiles) return;
  DISCO_NQ_DISPATCH(NQ, KLAUNCH(st, "k_mindt",
                                hipLaunchKernelGGL(k_mindt<Q>, dim3(L.n_itiles), dim3(TILE), 0, st, m, L.itiles)));
}

// plain copy of n doubles: the reference point for HBM byte counters
__global__ __launch_bounds__(TILE) void k_copy(double *dst, const double *src, long long n) {
  long long i = (long long)blockIdx.x * TILE + threadIdx.x;
  long long stride = (long long)gridDim.x * TILE;
  for (; i < n; i += stride) dst[i] = src[i];
}
void launch_copy(double *dst, const double *src, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_copy, dim3(2048), dim3(TILE), 0, st, dst, src, n);
}

void launch_faces(const DevMesh &m, const LaunchPlan &L, hipStream_t st, int NQ, int dir) {
  DISCO_NQ_DISPATCH(NQ, launch_faces_q<Q>(m, L, st, dir, 0));
}

template <int NQ>
static void launch_plm_q(const DevMesh &m, const LaunchPlan &L, hipStream_t st, int which) {
  if (which < 0) {
    hipLaunchKernelGGL((k_plm_p<NQ, false>), dim3(L.n_ztiles), dim3(TILE), 0, st, m, L.ztiles);
  } else {
    hipLaunchKernelGGL(k_plm_rz<NQ>, dim3(L.n_ztiles), dim3(TILE), 0, st, m, L.ztiles, which);
  }
}

void launch_plm(const DevMesh &m, const LaunchPlan &L, hipStream_t st, int NQ, int which) {
  DISCO_NQ_DISPATCH(NQ, launch_plm_q<Q>(m, L, st, which));
}

// cell_boundary_outflow_r on its own (the class API): both stages as two
// launches on the current state
void launch_bc_outflow_r(const DevMesh &m, const LaunchPlan &L, hipStream_t st, int NQ) {
  DISCO_NQ_DISPATCH(NQ, {
    if (L.n_bc0) hipLaunchKernelGGL(k_bc_r<Q>, dim3(L.n_bc0), dim3(TILE), 0, st, m, L.bc0);
    if (L.n_bc1) hipLaunchKernelGGL(k_bc_r<Q>, dim3(L.n_bc1), dim3(TILE), 0, st, m, L.bc1);
  });
}
