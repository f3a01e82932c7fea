// Unit-level kernels of the class API: the pieces timestep_substep calls
// (src/TimeStep/timestep_routines.c:16-113) as separate device launches, so
// that a driver can call cell_copy, cell_adjust_RK_cons, riemann_AddFlux,
// cell_add_src ... one by one as the reference does.  They work on the
// current state buffers plus a unit-level conserved state `uc` ([NQ][Ntot],
// the reference's cell->cons), and read the cells' tiph as stored, as the
// reference's routines read theCells[k][i][j].tiph.  One thread per cell,
// over the zone tiles; each cell's arithmetic is the reference's.
#include "cell_ops.h"
#include "unit.h"

namespace {

struct CellGeo {
  long long c;
  int i, k;
  double r, rp, rm, zp, zm, dz;
};
__device__ __forceinline__ bool cell_of(const DevMesh &m, const int2 tl, CellGeo &g) {
  const AnnInfo A = m.ann[tl.x];
  const int j = tl.y + threadIdx.x;
  if (j >= A.N) return false;
  g.c = A.zbase + j;
  g.i = A.i;
  g.k = A.k;
  g.rp = A.rp;
  g.rm = A.rm;
  g.zp = A.zp;
  g.zm = A.zm;
  g.r = .5 * (A.rp + A.rm);
  g.dz = A.zp - A.zm;
  return true;
}

// cell_calc_cons, src/Cell/cell_conversion.c:35-55
template <int NQ>
__global__ __launch_bounds__(TILE) void ku_derive(DevMesh m, const int2 *tiles, double *uc) {
  const int2 tl = tiles[blockIdx.x];
  CellGeo g;
  if (!cell_of(m, tl, g)) return;
  const DiskParams &P = m.dp[m.ann[tl.x].disk];
  double pr[NQ], u[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) pr[q] = m.prim[q * m.Ntot + g.c];
  d_prim2cons<NQ>(pr, u, g.r, .5 * (g.rp * g.rp - g.rm * g.rm) * m.dphi[g.c] * g.dz, P.gamma, P.NC);
#pragma unroll
  for (int q = 0; q < NQ; ++q) uc[q * m.Ntot + g.c] = u[q];
}

// cell_copy, src/Cell/cell_misc.c:28-41
template <int NQ>
__global__ __launch_bounds__(TILE) void ku_copy(DevMesh m, const int2 *tiles, const double *uc) {
  CellGeo g;
  if (!cell_of(m, tiles[blockIdx.x], g)) return;
#pragma unroll
  for (int q = 0; q < NQ; ++q) m.RKcons[q * m.Ntot + g.c] = uc[q * m.Ntot + g.c];
  m.RKtiph[g.c] = m.tiph[g.c];
}

// cell_adjust_RK_cons, src/Cell/cell_misc.c:43-57
template <int NQ>
__global__ __launch_bounds__(TILE) void ku_adjust(DevMesh m, const int2 *tiles, double *uc, double RK) {
  CellGeo g;
  if (!cell_of(m, tiles[blockIdx.x], g)) return;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const long long x = q * m.Ntot + g.c;
    uc[x] = (1.0 - RK) * uc[x] + RK * m.RKcons[x];
  }
}

// cell_add_src (part 0), cell_add_visc_src (1), cell_add_visc_src_old (2),
// cell_add_split_fictitious (3): src/Cell/cell_source.c:73-300,
// src/Cell/cell_add_split_fictitious.c:11-44.  Part 0 stores each tile's
// sink accretion of the two masses in mdot[2 tile + p] (host adds them in
// tile order) and each cooled cell's rate in cool.
template <int NQ>
__global__ __launch_bounds__(TILE) void ku_src(DevMesh m, const int2 *tiles, double *uc, double *cool, double *mdot,
                                               double dt, int part) {
  __shared__ double s_md[2][TILE / 64];
  const int2 tl = tiles[blockIdx.x];
  CellGeo g;
  double md0 = 0.0, md1 = 0.0;
  if (cell_of(m, tl, g)) {
    const AnnInfo A = m.ann[tl.x];
    const DiskParams &P = m.dp[A.disk];
    const DiskScalars &sc = m.sc[A.disk];
    const double4 rot = ann_rotation(m, P, tl.x);
    double pr[NQ], u[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      pr[q] = m.prim[q * m.Ntot + g.c];
      u[q] = uc[q * m.Ntot + g.c];
    }
    const double dphi = m.dphi[g.c];
    const double dV = dphi * .5 * (g.rp * g.rp - g.rm * g.rm) * g.dz;
    const double phi = m.tiph[g.c] - .5 * dphi;
    double cphi = 0.0, sphi = 0.0;
    if (part <= 1) sincos(phi, &sphi, &cphi);
    if (part == 0) {
      double cl = 0.0;
      src_grav_sink_cool<F_ALL>(P, sc, rot, g.i, g.k, g.r, dV, .5 * (g.zp + g.zm), cphi, sphi, pr, u, dt, &md0,
                                &md1, &cl);
      if (P.cooling == 1 && cool) cool[g.c] = cl;
    } else if (part == 1) {
      src_visc<F_ALL>(P, sc, rot, g.r, dV, cphi, sphi, pr, m.grad[UPP * m.Ntot + g.c], m.gradp[URR * m.Ntot + g.c],
                      u, dt);
    } else if (part == 2) {  // cell_add_visc_src_old, :279-300
      double nu = P.visc;
      u[SRR] += -dt * dV * nu * pr[RHO] * pr[URR] / g.r / g.r;
    } else {
      src_fict<F_ALL>(rot, g.r, dt, u);
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) uc[q * m.Ntot + g.c] = u[q];
  }
  if (part == 0 && mdot) {
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
      mdot[2 * blockIdx.x] = t0;
      mdot[2 * blockIdx.x + 1] = t1;
    }
  }
}

// cell_update_phi, src/Cell/cell_misc.c:59-79 (in place: each cell reads and
// writes only its own edge)
__global__ __launch_bounds__(TILE) void ku_update_phi(DevMesh m, const int2 *tiles, double RK, double dt) {
  const int2 tl = tiles[blockIdx.x];
  CellGeo g;
  if (!cell_of(m, tl, g)) return;
  const DiskParams &P = m.dp[m.ann[tl.x].disk];
  double t = m.tiph_new[g.c];
  double rk = m.RKtiph[g.c];
  int n = 0;
  while (t - rk > HALF_TURN && n++ < MAXWRAP) rk += TWO_PI;
  while (t - rk < -HALF_TURN && n++ < MAXWRAP) rk -= TWO_PI;
  m.RKtiph[g.c] = rk;
  t = (1.0 - RK) * t + RK * rk;
  const double w_total = m.wiph[tl.x] + ann_rotation(m, P, tl.x).x;
  t += w_total * dt / g.r;
  m.tiph_new[g.c] = t;
}

// cell_update_dphi, src/Cell/cell_misc.c:81-96 (reads edges, writes widths)
__global__ __launch_bounds__(TILE) void ku_update_dphi(DevMesh m, const int2 *tiles) {
  const int2 tl = tiles[blockIdx.x];
  const AnnInfo A = m.ann[tl.x];
  const int j = tl.y + threadIdx.x;
  if (j >= A.N) return;
  const int jm = j == 0 ? A.N - 1 : j - 1;
  m.dphi[A.zbase + j] = wrap_p(m.tiph[A.zbase + j] - m.tiph[A.zbase + jm]);
}

// cell_calc_prim, src/Cell/cell_conversion.c:105-122
template <int NQ>
__global__ __launch_bounds__(TILE) void ku_calc_prim(DevMesh m, const int2 *tiles, const double *uc) {
  const int2 tl = tiles[blockIdx.x];
  CellGeo g;
  if (!cell_of(m, tl, g)) return;
  const DiskParams &P = m.dp[m.ann[tl.x].disk];
  double u[NQ], pr[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) u[q] = uc[q * m.Ntot + g.c];
  d_cons2prim<NQ>(u, pr, g.r, .5 * (g.rp * g.rp - g.rm * g.rm) * m.dphi[g.c] * g.dz, P);
#pragma unroll
  for (int q = 0; q < NQ; ++q) m.prim[q * m.Ntot + g.c] = pr[q];
}

// riemann_setup_rz / riemann_setup_p + riemann_AddFlux for one face
// (src/Riemann/Riemann_setup.c:11-80, Riemann_routines.c:314-397)
template <int NQ>
__global__ __launch_bounds__(64) void ku_face(DevMesh m, double *uc, UnitFace f, double dt) {
  if (threadIdx.x != 0) return;
  const AnnInfo A = m.ann[f.ann_L];
  const DiskParams &P = m.dp[A.disk];
  const DiskScalars &sc = m.sc[A.disk];
  const long long cL = f.cL, cR = f.cR, N = m.Ntot;
  double primL[NQ], primR[NQ];
  double w = 0.0;
  int dir;
  if (f.kind == 1) {  // riemann_setup_p, :46-80
    const double dpL = m.dphi[cL], dpR = m.dphi[cR];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      primL[q] = m.prim[q * N + cL] + 0.5 * m.gradp[q * N + cL] * dpL;
      primR[q] = m.prim[q * N + cR] - 0.5 * m.gradp[q * N + cR] * dpR;
    }
    w = m.wiph[f.ann_L];  // cell_wiph(cL), Riemann_routines.c:323-324
    dir = 1;
  } else {  // riemann_setup_rz, :11-44
    const double pL = m.tiph[cL] - .5 * m.dphi[cL];
    const double pR = m.tiph[cR] - .5 * m.dphi[cR];
    const double dpL = wrap_pm(f.cm - pL);
    const double dpR = wrap_pm(-f.cm + pR);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      primL[q] = m.prim[q * N + cL] + m.grad[q * N + cL] * f.deltaL + m.gradp[q * N + cL] * dpL;
      primR[q] = m.prim[q * N + cR] - m.grad[q * N + cR] * f.deltaR - m.gradp[q * N + cR] * dpR;
    }
    dir = f.kind == 0 ? 0 : 2;
  }
  ViscIn vi;
  if (P.visc > 0.0) {
    vi.prim = m.prim;
    vi.grad = m.grad;
    vi.gradp = m.gradp;
    vi.stride = N;
    vi.cL = cL;
    vi.cR = cR;
    vi.tL = m.tiph[cL];
    vi.tR = m.tiph[cR];
    vi.dL = m.dphi[cL];
    vi.dR = m.dphi[cR];
    vi.cm = f.kind == 1 ? m.tiph[cL] : f.cm;  // riemann_setup_p: cm = cell_tiph(cL), Riemann_setup.c:69
    vi.r_cell_L = f.r_cell_L;
    vi.r_cell_R = f.r_cell_R;
    const double4 rot = ann_rotation(m, P, f.ann_L);
    vi.rdr_om = f.kind == 0 ? rot.w : rot.y;
  }
  double X[NQ], Y[3] = {0.0, 0.0, 0.0};
  const int e = d_face_flux<NQ>(primL, primR, dir, f.r, f.dA, w, dt, P, sc, &vi, X, Y);
  if (e) {
    set_err(m.err, e);
    return;
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    uc[q * N + cL] += -X[q];
    uc[q * N + cR] += X[q];
  }
  if (P.visc > 0.) {
    uc[SRR * N + cL] += -Y[0] / f.r_cell_L;
    uc[SRR * N + cR] += Y[0] / f.r_cell_R;
    uc[LLL * N + cL] += -Y[1];
    uc[LLL * N + cR] += Y[1];
    uc[TAU * N + cL] += -Y[2];
    uc[TAU * N + cR] += Y[2];
  }
}

// cell_add_cons, src/Cell/cell_modify_data.c:11-13
__global__ void ku_add(double *uc, long long x, double v) {
  if (threadIdx.x == 0) uc[x] += v;
}

// cell_boundary_outflow_z, src/Cell/cell_boundary.c:141-206: layer 0 takes
// the overlap-area-weighted primitives of layer 1 (its z faces as L), layer
// NZ-1 those of layer NZ-2 (its faces as R); both flip v_z
template <int NQ>
__global__ __launch_bounds__(TILE) void ku_outflow_z(DevMesh m, const int2 *tiles) {
  const int2 tl = tiles[blockIdx.x];
  const AnnInfo A = m.ann[tl.x];
  const DiskParams &P = m.dp[A.disk];
  const int j = tl.y + threadIdx.x;
  if (j >= A.N || (A.k != 0 && A.k != P.NZ - 1)) return;
  const long long c = A.zbase + j;
  double v[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) v[q] = 0.0;
  if (A.k == 0) {
    faces_as_L(m, zpair_of(P, 0, A.i), m.fo_z, c, j, A.N, [&](const PairInfo &PI, long long g) {
      const int cR = m.fR[g];
      const double dA = PI.r * m.fdphi[g] * PI.deltaPerp;
#pragma unroll
      for (int q = 0; q < NQ; ++q) v[q] += m.prim[q * m.Ntot + cR] * dA;
    });
  } else {
    faces_as_R(m, zpair_of(P, P.NZ - 2, A.i), j, A.N, [&](const PairInfo &PI, long long g) {
      const int cL = m.fL[g];
      const double dA = PI.r * m.fdphi[g] * PI.deltaPerp;
#pragma unroll
      for (int q = 0; q < NQ; ++q) v[q] += m.prim[q * m.Ntot + cL] * dA;
    });
  }
  const double dA = .5 * (A.rp * A.rp - A.rm * A.rm) * m.dphi[c];
#pragma unroll
  for (int q = 0; q < NQ; ++q) v[q] /= dA;
  v[UZZ] *= -1.0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) m.prim[q * m.Ntot + c] = v[q];
}

// cell_boundary_fixed_r (mode UNIT_FIXED_R), cell_boundary_fixed_z
// (UNIT_FIXED_Z), cell_bc_damp (UNIT_DAMP): src/Cell/cell_boundary.c:208-254,
// src/Cell/cell_misc.c:98-132, from the plugin's rows
template <int NQ>
__global__ __launch_bounds__(TILE) void ku_bc_init(DevMesh m, const int2 *tiles, const InitRow *rows, int mode,
                                                   double dt) {
  const int2 tl = tiles[blockIdx.x];
  const AnnInfo A = m.ann[tl.x];
  const DiskParams &P = m.dp[A.disk];
  const int j = tl.y + threadIdx.x;
  if (j >= A.N) return;
  const InitRow R = rows[tl.x];
  bool act;
  if (mode == UNIT_FIXED_R)
    act = (A.i < P.NgMinR && P.no_inner_bc != 1) || A.i >= P.NR - P.NgMaxR;
  else if (mode == UNIT_FIXED_Z)
    act = A.k < P.NgMinZ || A.k >= P.NZ - P.NgMaxZ;
  else
    act = (R.flags & INIT_DAMP) != 0;
  if (!act) return;
  const long long c = A.zbase + j;
  double pr[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) pr[q] = m.prim[q * m.Ntot + c];
  if (mode != UNIT_DAMP) {
    single_init<NQ>(P.initial_data_type, R, .5 * (A.rm + A.rp), m.tiph[c], m.dphi[c], pr);
  } else {
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const double dprim = R.prim[q] - pr[q];
      pr[q] += (1. - exp(-R.rate * dt)) * dprim;
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) m.prim[q * m.Ntot + c] = pr[q];
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers: every zone tile of the batch (table order)
// ---------------------------------------------------------------------------
void launch_unit(const DevMesh &m, const LaunchPlan &L, hipStream_t st, int NQ, int op, double *uc, double a,
                 double b, double *cool, double *mdot, int arg) {
  const dim3 G(L.n_ztiles), B(TILE);
  const int2 *tl = L.ztiles;
  DevMesh m2 = m;  // cell_update_phi reads and writes the current edges
  m2.tiph_new = const_cast<double *>(m.tiph);
  DISCO_NQ_DISPATCH(NQ, {
    switch (op) {
      case UNIT_DERIVE: hipLaunchKernelGGL(ku_derive<Q>, G, B, 0, st, m, tl, uc); break;
      case UNIT_COPY: hipLaunchKernelGGL(ku_copy<Q>, G, B, 0, st, m, tl, uc); break;
      case UNIT_ADJUST: hipLaunchKernelGGL(ku_adjust<Q>, G, B, 0, st, m, tl, uc, a); break;
      case UNIT_SRC: hipLaunchKernelGGL(ku_src<Q>, G, B, 0, st, m, tl, uc, cool, mdot, a, arg); break;
      case UNIT_UPDATE_PHI: hipLaunchKernelGGL(ku_update_phi, G, B, 0, st, m2, tl, a, b); break;
      case UNIT_UPDATE_DPHI: hipLaunchKernelGGL(ku_update_dphi, G, B, 0, st, m, tl); break;
      case UNIT_CALC_PRIM: hipLaunchKernelGGL(ku_calc_prim<Q>, G, B, 0, st, m, tl, uc); break;
      case UNIT_OUTFLOW_Z: hipLaunchKernelGGL(ku_outflow_z<Q>, G, B, 0, st, m, tl); break;
      default: break;
    }
  });
}

void launch_unit_bc(const DevMesh &m, const LaunchPlan &L, hipStream_t st, int NQ, const InitRow *rows, int mode,
                    double dt) {
  DISCO_NQ_DISPATCH(NQ, hipLaunchKernelGGL(ku_bc_init<Q>, dim3(L.n_ztiles), dim3(TILE), 0, st, m, L.ztiles, rows,
                                           mode, dt));
}

void launch_unit_face(const DevMesh &m, hipStream_t st, int NQ, double *uc, const UnitFace &f, double dt) {
  DISCO_NQ_DISPATCH(NQ, hipLaunchKernelGGL(ku_face<Q>, dim3(1), dim3(64), 0, st, m, uc, f, dt));
}

void launch_unit_add(double *uc, long long x, double v, hipStream_t st) {
  hipLaunchKernelGGL(ku_add, dim3(1), dim3(64), 0, st, uc, x, v);
}
