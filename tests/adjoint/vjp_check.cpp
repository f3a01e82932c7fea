// Central-difference checks of the hand-written VJPs (disco2_amd/csrc/adjoint.h)
// against the forward functions they differentiate (physics.h, the same code
// the kernels run), on random states.  Host only: built with
// hipcc --offload-host-only.  Prints one line per check:
//   name trials failures worst_relative_error
#include <cmath>
#include <cstring>
#include <cstdio>
#include <random>
#include <vector>

#include "../../disco2_amd/csrc/adjoint.h"

struct Stat {
  const char *name;
  int n = 0, fail = 0;
  double worst = 0.0;
  void print() const { printf("%s %d %d %.3e\n", name, n, fail, worst); }
};

static std::mt19937_64 rng(12345);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }

// Compare g = J^T yb (from vjp) with central differences of <yb, f(x)>.
template <class F, class G>
static void check(Stat &st, const std::vector<double> &x, int nout, F f, G vjp, double tol = 1e-6) {
  const int nin = (int)x.size();
  std::vector<double> yb(nout), g(nin, 0.0);
  for (auto &v : yb) v = U(-1, 1);
  vjp(x.data(), yb.data(), g.data());
  std::vector<double> fd(nin), yp(nout), ym(nout);
  double gmax = 0.0;
  for (int i = 0; i < nin; ++i) {
    const double h = 1e-6 * fmax(fabs(x[i]), 1e-2);
    std::vector<double> xp = x, xm = x;
    xp[i] += h;
    xm[i] -= h;
    f(xp.data(), yp.data());
    f(xm.data(), ym.data());
    double s = 0.0;
    for (int k = 0; k < nout; ++k) s += yb[k] * (yp[k] - ym[k]);
    fd[i] = s / (2. * h);
    gmax = fmax(gmax, fabs(fd[i]));
  }
  double err = 0.0;
  for (int i = 0; i < nin; ++i) err = fmax(err, fabs(g[i] - fd[i]));
  err /= fmax(gmax, 1e-300);
  st.n++;
  if (!(err <= tol)) st.fail++;
  st.worst = fmax(st.worst, err);
}

static DiskParams params(int riemann, double visc) {
  DiskParams P = {};
  P.gamma = 5. / 3.;
  P.cfl = 0.4;
  P.plm = 1.0;
  P.visc = visc;
  P.visc_const = 0;
  P.initial_data_type = 1;
  P.g_eps = 0.05;
  P.rho_floor = 1e-10;
  P.cs_floor = 1e-10;
  P.cs_cap = 1e10;
  P.vel_cap = 1e10;
  P.riemann = riemann;
  P.grav2d = 1;
  P.phi_order = 2;
  P.NC = 5;
  return P;
}
static DiskScalars masses(const double *m) {  // m: M0, M1, xbh0, ybh0, xbh1, ybh1
  DiskScalars sc = {};
  sc.gm[0].M = m[0];
  sc.gm[1].M = m[1];
  sc.xbh[0] = m[2];
  sc.ybh[0] = m[3];
  sc.xbh[1] = m[4];
  sc.ybh[1] = m[5];
  sc.epsn = 0.05 * 0.05;
  return sc;
}
static void rand_prim(double *p) {
  p[RHO] = U(0.5, 2.0);
  p[PPP] = U(0.01, 0.5);
  p[URR] = U(-0.3, 0.3);
  p[UPP] = U(0.2, 1.2);
  p[UZZ] = 0.0;
}
static void rand_masses(double *m) {
  m[0] = U(0.3, 0.7);
  m[1] = 1.0 - m[0];
  double ph = U(0, 6.28);
  m[2] = m[1] * cos(ph);
  m[3] = m[1] * sin(ph);
  m[4] = -m[0] * cos(ph);
  m[5] = -m[0] * sin(ph);
}
static void add_scalars(const ScalarAdj &sb, double *g) {  // M0 M1 x0 y0 x1 y1
  g[0] += sb.M[0];
  g[1] += sb.M[1];
  g[2] += sb.xbh[0];
  g[3] += sb.ybh[0];
  g[4] += sb.xbh[1];
  g[5] += sb.ybh[1];
}

int main() {
  const int T = 300;
  {  // prim2cons
    Stat st{"prim2cons"};
    for (int t = 0; t < T; ++t) {
      std::vector<double> x(5);
      rand_prim(x.data());
      x[UZZ] = U(-0.1, 0.1);
      const double r = U(0.5, 3.0), dV = U(0.01, 0.2), G = 5. / 3.;
      check(st, x, 5, [&](const double *p, double *y) { d_prim2cons<5>(p, y, r, dV, G, 5); },
            [&](const double *p, const double *yb, double *g) { vjp_prim2cons(p, r, dV, G, yb, g); });
    }
    st.print();
  }
  for (int mode = 0; mode < 3; ++mode) {  // cons2prim: plain, pressure floor, velocity cap
    Stat st{mode == 0 ? "cons2prim" : (mode == 1 ? "cons2prim_pfloor" : "cons2prim_vcap")};
    for (int t = 0; t < T; ++t) {
      double p[5];
      rand_prim(p);
      p[UZZ] = U(-0.1, 0.1);
      const double r = U(0.5, 3.0), dV = U(0.01, 0.2);
      DiskParams P = params(0, -1);
      if (mode == 1) {
        p[PPP] = 1e-6;
        P.cs_floor = 0.01;
      }
      if (mode == 2) P.vel_cap = 0.3;
      std::vector<double> x(5);
      d_prim2cons<5>(p, x.data(), r, dV, P.gamma, 5);
      check(st, x, 5, [&](const double *c, double *y) { d_cons2prim<5>(c, y, r, dV, P); },
            [&](const double *c, const double *yb, double *g) { vjp_cons2prim(c, r, dV, P, yb, g); });
    }
    st.print();
  }
  for (int solver = 0; solver < 2; ++solver)
    for (int dir = 0; dir < 3; ++dir) {  // the Riemann flux in each of the four states, dt as an input
      char nm[64];
      snprintf(nm, sizeof nm, "%s_dir%d", solver == 0 ? "hllc" : "hll", dir);
      Stat st{strdup(nm)};
      for (int t = 0; t < T; ++t) {
        DiskParams P = params(solver == 0 ? HLLC_SOLVER : HLL_SOLVER, -1.0);
        std::vector<double> x(11);
        rand_prim(&x[0]);
        rand_prim(&x[5]);
        x[UZZ] = U(-0.2, 0.2);
        x[5 + UZZ] = U(-0.2, 0.2);
        x[10] = U(1e-3, 1e-2);
        const double r = U(0.5, 3.0), dA = U(0.01, 0.1);
        // face velocity in each of the four wave-speed intervals in turn
        const int n0 = dir == 0, n1 = dir == 1, n2 = dir == 2;
        const double vnL = x[URR] * n0 + r * x[UPP] * n1 + x[UZZ] * n2;
        const double vnR = x[5 + URR] * n0 + r * x[5 + UPP] * n1 + x[5 + UZZ] * n2;
        const double cL = sqrt(P.gamma * x[PPP] / x[RHO]), cR = sqrt(P.gamma * x[5 + PPP] / x[5 + RHO]);
        const double Sl = fmin(vnL - cL, vnR - cR), Sr = fmax(vnL + cL, vnR + cR);
        const double Ss = (x[5 + RHO] * vnR * (Sr - vnR) - x[RHO] * vnL * (Sl - vnL) + x[PPP] - x[5 + PPP]) /
                          (x[5 + RHO] * (Sr - vnR) - x[RHO] * (Sl - vnL));
        const double ws[4] = {Sl - 0.5, 0.5 * (Sl + Ss), 0.5 * (Ss + Sr), Sr + 0.5};
        const double w = ws[t % 4];
        DiskScalars sc = {};
        auto f = [&](const double *v, double *y) {
          double Y[3];
          d_face_flux<5>(v, v + 5, dir, r, dA, w, v[10], P, sc, nullptr, y, Y);
        };
        auto g = [&](const double *v, const double *yb, double *gx) {
          ScalarAdj sb = {};
          double Yb[3] = {0, 0, 0};
          vjp_face_flux(P, sc, v, v + 5, dir, r, dA, w, v[10], 0.0, 0.0, nullptr, nullptr, nullptr, yb, Yb, gx,
                        gx + 5, nullptr, nullptr, nullptr, sb);
          gx[10] += sb.dt;
        };
        check(st, x, 5, f, g);
      }
      st.print();
    }
  for (int kind = 0; kind < 2; ++kind) {  // viscous r and phi faces: cells, dt, alpha, masses
    Stat st{kind == 0 ? "rface_visc" : "phiface_visc"};
    for (int t = 0; t < T; ++t) {
      const int solver = t % 2;
      // x: pL pR gL gR gpL gpR (30) | dt alpha | M0 M1 x0 y0 x1 y1 | geometry:
      // r face: dA cm tL tR; phi face: a rigid rotation of the annulus
      std::vector<double> x(kind == 0 ? 42 : 39);
      rand_prim(&x[0]);
      rand_prim(&x[5]);
      for (int q = 0; q < 5; ++q) {  // no v_z anywhere: the viscous flux refuses it
        const bool z = q == UZZ;
        x[10 + q] = z ? 0.0 : U(-0.05, 0.05);
        x[15 + q] = z ? 0.0 : U(-0.05, 0.05);
        x[20 + q] = z ? 0.0 : U(-0.05, 0.05);
        x[25 + q] = z ? 0.0 : U(-0.05, 0.05);
      }
      x[30] = U(1e-3, 1e-2);
      x[31] = U(1e-3, 1e-1);
      rand_masses(&x[32]);
      RFaceGeo rg;
      rg.r = U(1.0, 3.0);
      rg.deltaL = U(0.01, 0.05);
      rg.deltaR = U(0.01, 0.05);
      rg.dA = U(0.001, 0.01);
      rg.tL = U(0.2, 6.0);
      rg.dL = U(0.01, 0.05);
      rg.tR = rg.tL + U(-0.02, 0.02);
      rg.dR = U(0.01, 0.05);
      rg.cm = rg.tL - U(0.0, rg.dL);
      rg.rdr = 0.0;
      PhiFaceGeo pg;
      pg.r = U(1.0, 3.0);
      pg.dA = U(0.001, 0.01);
      pg.w = U(0.3, 1.0);
      pg.tL = U(0.2, 6.0);
      pg.dL = U(0.01, 0.05);
      pg.dR = U(0.01, 0.05);
      pg.tR = pg.tL + pg.dR;
      pg.rdr = 0.0;
      if (kind == 0) {
        x[38] = rg.dA;
        x[39] = rg.cm;
        x[40] = rg.tL;
        x[41] = rg.tR;
      } else {
        x[38] = 0.0;
      }
      auto unpack = [&](const double *v, FacePair &c, DiskParams &P, DiskScalars &sc) {
        for (int q = 0; q < 5; ++q) {
          c.pL[q] = v[q];
          c.pR[q] = v[5 + q];
          c.gL[q] = v[10 + q];
          c.gR[q] = v[15 + q];
          c.gpL[q] = v[20 + q];
          c.gpR[q] = v[25 + q];
        }
        P = params(solver == 0 ? HLLC_SOLVER : HLL_SOLVER, v[31]);
        sc = masses(v + 32);
      };
      auto f = [&](const double *v, double *y) {
        FacePair c;
        DiskParams P;
        DiskScalars sc;
        unpack(v, c, P, sc);
        if (kind == 0) {
          RFaceGeo g = rg;
          g.dA = v[38];
          g.cm = v[39];
          g.tL = v[40];
          g.tR = v[41];
          rface_flux(P, sc, c, g, v[30], y, y + 5);
        } else {
          PhiFaceGeo g = pg;
          g.tL += v[38];
          g.tR += v[38];
          phiface_flux(P, sc, c, g, v[30], y, y + 5);
        }
      };
      auto g = [&](const double *v, const double *yb, double *gx) {
        FacePair c, cb = {};
        DiskParams P;
        DiskScalars sc;
        unpack(v, c, P, sc);
        ScalarAdj sb = {};
        if (kind == 0) {
          RFaceGeo g = rg, gb = {};
          g.dA = v[38];
          g.cm = v[39];
          g.tL = v[40];
          g.tR = v[41];
          vjp_rface_flux(P, sc, c, g, v[30], yb, yb + 5, cb, sb, gb);
          gx[38] += gb.dA;
          gx[39] += gb.cm;
          gx[40] += gb.tL;
          gx[41] += gb.tR;
        } else {
          PhiFaceGeo g = pg;
          g.tL += v[38];
          g.tR += v[38];
          double tb = 0.0;
          vjp_phiface_flux(P, sc, c, g, v[30], yb, yb + 5, cb, sb, &tb);
          gx[38] += tb;
        }
        for (int q = 0; q < 5; ++q) {
          gx[q] += cb.pL[q];
          gx[5 + q] += cb.pR[q];
          gx[10 + q] += cb.gL[q];
          gx[15 + q] += cb.gR[q];
          gx[20 + q] += cb.gpL[q];
          gx[25 + q] += cb.gpR[q];
        }
        gx[30] += sb.dt;
        gx[31] += sb.alpha;
        add_scalars(sb, gx + 32);
      };
      check(st, x, 8, f, g);
    }
    st.print();
  }
  for (int g2d = 1; g2d >= 0; --g2d) {  // sources and the torque term: prim, dt, masses
    Stat st{g2d ? "sources_2d" : "sources_3d"};
    for (int t = 0; t < T; ++t) {
      std::vector<double> x(13);
      rand_prim(&x[0]);
      x[UZZ] = U(-0.1, 0.1);
      x[5] = U(1e-3, 1e-2);
      rand_masses(&x[6]);
      x[12] = U(0, 6.28);  // the cell-centre angle
      const double r = U(0.6, 3.0), dV = U(0.001, 0.05), zc = g2d ? 0.0 : U(-0.3, 0.3);
      DiskParams P = params(0, -1);
      P.grav2d = g2d;
      auto f = [&](const double *v, double *y) {
        DiskScalars sc = masses(v + 6);
        double u[5] = {0, 0, 0, 0, 0}, Fp, s, c;
        sincos(v[12], &s, &c);
        add_geom_grav_src(u, v, r, dV, v[5], c, s, zc, 0.0, 0.0, 0.0, P, sc, &Fp);
        for (int q = 0; q < 5; ++q) y[q] = u[q];
        y[5] = v[5] * dV * v[RHO] * Fp * r;
      };
      auto g = [&](const double *v, const double *yb, double *gx) {
        DiskScalars sc = masses(v + 6);
        ScalarAdj sb = {};
        vjp_cell_sources(P, sc, v, r, dV, v[5], v[12], zc, yb, yb[5], gx, sb, &gx[12]);
        gx[5] += sb.dt;
        add_scalars(sb, gx + 6);
      };
      check(st, x, 6, f, g);
    }
    st.print();
  }
  for (int vmode = 0; vmode < 3; ++vmode) {  // CFL dt: no viscosity, constant, alpha law
    Stat st{vmode == 0 ? "cell_dt" : (vmode == 1 ? "cell_dt_nu_const" : "cell_dt_alpha")};
    for (int t = 0; t < T; ++t) {
      std::vector<double> x(13);
      rand_prim(&x[0]);
      x[UZZ] = U(-0.1, 0.1);
      x[5] = U(1e-3, 1e-1);
      rand_masses(&x[6]);
      x[12] = U(0, 6.28);  // the cell-centre angle
      const double r = U(0.6, 3.0), dx = U(0.01, 0.05), w = U(0.0, 1.0);
      auto mkP = [&](double alpha) {
        DiskParams P = params(0, vmode == 0 ? -1.0 : alpha);
        P.visc_const = vmode == 1;
        return P;
      };
      auto f = [&](const double *v, double *y) {
        DiskParams P = mkP(v[5]);
        DiskScalars sc = masses(v + 6);
        y[0] = cell_dt_value(P, sc, v, r, dx, w, v[12]);
      };
      auto g = [&](const double *v, const double *yb, double *gx) {
        DiskParams P = mkP(v[5]);
        DiskScalars sc = masses(v + 6);
        ScalarAdj sb = {};
        vjp_cell_dt(P, sc, v, r, dx, w, v[12], yb[0], gx, sb, &gx[12]);
        gx[5] += sb.alpha;
        add_scalars(sb, gx + 6);
      };
      check(st, x, 1, f, g);
    }
    st.print();
  }
  {  // phi limiter (cell_plm.c:141-153): weights against differences of the limited slope
    Stat st{"plm_p"};
    for (int t = 0; t < T; ++t) {
      std::vector<double> x = {U(0, 1), U(0, 1), U(0, 1)};
      const double dpL = U(0.01, 0.05), dpR = U(0.01, 0.05), PLM = t % 2 ? 1.0 : 2.0;
      auto f = [&](const double *v, double *y) {
        double sL = (v[1] - v[0]) / dpL, sR = (v[2] - v[1]) / dpR, sM = (v[2] - v[0]) / (dpL + dpR);
        double s = PLM * sL;
        if (fabs(s) > PLM * fabs(sR)) s = PLM * sR;
        if (fabs(s) > fabs(sM)) s = sM;
        if (sL * sR < 0.0) s = 0.0;
        y[0] = s;
      };
      auto g = [&](const double *v, const double *yb, double *gx) {
        double wL, wc, wR;
        plm_p_weights(v[0], v[1], v[2], dpL, dpR, PLM, &wL, &wc, &wR);
        gx[0] += wL * yb[0];
        gx[1] += wc * yb[0];
        gx[2] += wR * yb[0];
      };
      check(st, x, 1, f, g);
    }
    st.print();
  }
  return 0;
}
