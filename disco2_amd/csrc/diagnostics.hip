// Device side of the diagnostics (src/Diagnostics/diagnostics_set.c:13-303):
// one measurement of one disk.
//
// k_diag_ann runs one workgroup per interior annulus i.  Its threads form the
// per-cell terms of a chunk of cells in LDS (the phi-averaged profile terms
// x / N_p * dz, and rho dV under each mass-count condition) and write the
// equatorial cells; then one thread per quantity adds the chunk's terms in
// cell order.  Over the layers k and the cells j of annulus i that is the
// reference's own order (its loops run k, i, j and each profile sum belongs
// to one i), so the profiles are bitwise the reference's.  The masses are
// summed the same way per annulus; the reference keeps one running sum over
// the whole disk, so they agree to rounding.  k_diag_fin forms the scalars
// (sequential over annuli, as the reference), the inner-wedge mass, the Mdot
// probe and the time-step scalings.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "diagnostics.h"

namespace {

constexpr int NV = DIAG_NV, NE = DIAG_NE, NS = DIAG_NS, NM = 5;
constexpr int NT = 12 + NM;  // terms per cell: 12 profile terms + 5 mass terms

__device__ __forceinline__ double grav_dist0(double rb, double phib, double r, double phi) {
  // gravMass_dist, src/GravMass/GravMass_access.c:25-31, at z = 0
  const double xbh = rb * cos(phib), ybh = rb * sin(phib);
  const double x = r * cos(phi), y = r * sin(phi);
  return sqrt((x - xbh) * (x - xbh) + (y - ybh) * (y - ybh) + 0. * 0.);
}

__global__ __launch_bounds__(TILE) void k_diag_ann(DevMesh m, DiagArgs g) {
  const int ii = blockIdx.x, i = g.imin + ii, t = threadIdx.x;
  const DiskParams &P = m.dp[g.disk];
  const DiskScalars &sc = m.sc[g.disk];
  __shared__ double s[NT][TILE];
  double acc = 0.0;
  for (int k = g.kmin; k < g.kmax; ++k) {
    const AnnInfo A = m.ann[P.ann_base + k * P.NR + i];
    const double zp = A.zp, zm = A.zm;
    const double z = 0.5 * (zm + zp), dz = zp - zm;
    const double rp = A.rp, rm = A.rm;
    const double r = 0.5 * (rm + rp);
    const double a = sc.gm[0].r + sc.gm[1].r;
    const bool equat = (fabs(zp) < 0.0000001) || (fabs(z) < 0.0000001);
    const double np = A.N;
    for (int j0 = 0; j0 < A.N; j0 += TILE) {
      const int j = j0 + t;
      if (j < A.N) {
        const long long c = A.zbase + j;
        const double dphi = m.dphi[c];
        const double phi = m.tiph[c] - 0.5 * dphi;
        const double rho = m.prim[RHO * m.Ntot + c];
        const double press = m.prim[PPP * m.Ntot + c];
        const double vr = m.prim[URR * m.Ntot + c];
        const double vp_minus_w = m.prim[UPP * m.Ntot + c] * r;
        const double vp = m.prim[UPP * m.Ntot + c] * r + d_rOm_a(P.w_analytic, r, a);
        const double vz = m.prim[UZZ * m.Ntot + c];
        const double v2 = vr * vr + vp * vp + vz * vz;
        const double rhoe = press / (P.gamma - 1.);
        const double KE = 0.5 * rho * v2;
        const double E_hydro = rhoe + KE;
        const double Cool = 0.0;  // no cooling (checked on the host)
        const double dV = 0.5 * (rp * rp - rm * rm) * dphi;
        // the reference reads prim[5] whatever NUM_Q is; past the array at 5
        const double passive = P.NQ > 5 ? m.prim[5 * m.Ntot + c] : 0.0;
        const double d0 = grav_dist0(sc.gm[0].r, sc.gm[0].phi, r, phi);
        const double d1 = grav_dist0(sc.gm[1].r, sc.gm[1].phi, r, phi);
        const double dM = rho * dV;
        const double x[12] = {r, rho, press, vr, vp, vz, vp_minus_w, Cool, passive, E_hydro, KE, rhoe};
#pragma unroll
        for (int n = 0; n < 12; ++n) s[n][t] = (x[n] / np * dz);
        s[12][t] = rp <= 1.0 ? dM : 0.0;
        s[13][t] = d0 < 0.2 ? dM : 0.0;
        s[14][t] = d1 < 0.2 ? dM : 0.0;
        s[15][t] = d0 < 0.4 ? dM : 0.0;
        s[16][t] = d1 < 0.4 ? dM : 0.0;
        if (equat && g.eq_base[k] >= 0) {
          double *e = g.equat + (size_t)(g.eq_base[k] + g.ioff[ii] + j) * NE;
          const double v[13] = {r, phi, rho, press, vr, vp, vz, vp_minus_w, Cool, passive, E_hydro, KE, rhoe};
#pragma unroll
          for (int n = 0; n < 13; ++n) e[n] = v[n];
#pragma unroll
          for (int n = 13; n < NE; ++n) e[n] = 0.0;
        }
      }
      __syncthreads();
      if (t < NT) {
        const int n = min(TILE, A.N - j0);
#pragma unroll 16
        for (int cc = 0; cc < n; ++cc) acc += s[t][cc];
      }
      __syncthreads();
    }
  }
  if (t < 12) g.vec[(size_t)ii * NV + t] = acc;
  if (t >= 12 && t < NV) g.vec[(size_t)ii * NV + t] = 0.0;  // the placeholders: 0.0 + 0.0
  if (t >= 12 && t < NT) g.mpart[(size_t)ii * NM + (t - 12)] = acc;
}

// scalars, masses, inner wedge, Mdot probe, scalings (diagnostics_set.c:193-245).
// Every sum runs in the reference's order; the workgroup stages each chunk
// of its terms in LDS and one thread per sum adds them in order.
constexpr int FT = 256;
__global__ __launch_bounds__(FT) void k_diag_fin(DevMesh m, DiagArgs g) {
  const int t = threadIdx.x;
  const DiskParams &P = m.dp[g.disk];
  const int nrg = g.nrg;
  const double RMIN = g.rz[0], RMAX = g.rz[1], ZMIN = g.rz[2], ZMAX = g.rz[3];
  __shared__ double s[NS + NM][FT];
  double acc = 0.0;  // thread t < NS: scalar t; NS <= t < NS + NM: mass sum t - NS
  // ScalarDiag_temp[n] += VectorDiag_temp[i][n+1] * (rp*rp - rm*rm), i in order
  for (int c0 = 0; c0 < nrg; c0 += FT) {
    const int ii = c0 + t;
    if (ii < nrg) {
      const AnnInfo A = m.ann[P.ann_base + g.kmin * P.NR + g.imin + ii];
      const double area = A.rp * A.rp - A.rm * A.rm;
#pragma unroll
      for (int n = 0; n < NS; ++n) s[n][t] = g.vec[(size_t)ii * NV + n + 1] * area;
#pragma unroll
      for (int n = 0; n < NM; ++n) s[NS + n][t] = g.mpart[(size_t)ii * NM + n];
    }
    __syncthreads();
    if (t < NS + NM) {
      const int n = min(FT, nrg - c0);
#pragma unroll 16
      for (int cc = 0; cc < n; ++cc) acc += s[t][cc];
    }
    __syncthreads();
  }
  if (t < NS) g.scal[t] = acc * (g.dtout / ((ZMAX - ZMIN) * (RMAX * RMAX - RMIN * RMIN)));
  if (t >= NS && t < NS + NM) g.out_mass[t - NS] = acc;
  // inner wedge: cell (0, j, 0) for j below N_p of the annulus the previous
  // loop left (imax), i.e. the first cells of layer 0, beyond annulus 0 when
  // j >= N_p[0]
  {
    const AnnInfo A0 = m.ann[P.ann_base];
    const double rp = A0.rp, rm = A0.rm;
    const int n = m.ann[P.ann_base + g.imin + nrg].N;
    double w = 0.0;
    for (int c0 = 0; c0 < n; c0 += FT) {
      const int j = c0 + t;
      if (j < n) {
        const long long c = A0.zbase + j;
        const double dV = 0.5 * (rp * rp - rm * rm) * m.dphi[c];
        s[0][t] = m.prim[RHO * m.Ntot + c] * dV;
      }
      __syncthreads();
      if (t == 0) {
        const int nn = min(FT, n - c0);
#pragma unroll 16
        for (int cc = 0; cc < nn; ++cc) w += s[0][cc];
      }
      __syncthreads();
    }
    if (t == 0) g.out_mass[NM] = w;
  }
  // the Mdot probe: the first annulus with r > 1 (profiles not yet scaled)
  if (t == 0) {
    int found = 0;
    double Mdot = 0.0, rr = 0.0;
    for (int ii = 0; ii < nrg && !found; ++ii) {
      const double r = g.vec[(size_t)ii * NV] / (ZMAX - ZMIN);
      if (r > 1.0) {
        Mdot = g.vec[(size_t)ii * NV + 5] * 2. * M_PI * r / (ZMAX - ZMIN);
        rr = r;
        found = 1;
      }
    }
    g.out_mass[NM + 1] = Mdot;
    g.out_mass[NM + 2] = rr;
  }
  __syncthreads();
  const double f = g.dtout / (ZMAX - ZMIN);
  for (int x = t; x < nrg * NV; x += FT) g.vec[x] = g.vec[x] * f;
}

}  // namespace

void launch_diag(const DevMesh &m, const DiagArgs &g, hipStream_t st) {
  hipLaunchKernelGGL(k_diag_ann, dim3(g.nrg), dim3(TILE), 0, st, m, g);
  hipLaunchKernelGGL(k_diag_fin, dim3(1), dim3(FT), 0, st, m, g);
}
