// Initial state of a problem: mesh coordinates, boundary conditions and the
// test-case initialisers of vnr_remap/Init.cc:9-511 (plus the cell-mass,
// time-step, energy and pseudo-viscosity initialisation of
// PhaseLagrange.cc:14-24 and Init.cc:435-489), computed once on the host.
#include "init.hpp"

#include <cmath>
#include <cstring>
#include <limits>

namespace mhy {

namespace {

inline double smin(double a, double b) { return (b < a) ? b : a; }
inline double smax(double a, double b) { return (a < b) ? b : a; }

struct P2 {
  double x, y;
};
inline P2 add(P2 a, P2 b) { return {a.x + b.x, a.y + b.y}; }
inline P2 sub(P2 a, P2 b) { return {a.x - b.x, a.y - b.y}; }
inline P2 mul(P2 a, double s) { return {a.x * s, a.y * s}; }
inline double cross(P2 a, P2 b) { return a.x * b.y - a.y * b.x; }
inline double norm(P2 a) {
  double r = 0.0;
  r = r + a.x * a.x;
  r = r + a.y * a.y;
  return std::sqrt(r);
}

// splitmix64, the random stream of the randomised initialisers
struct Rng {
  unsigned long long s;
  explicit Rng(unsigned long long seed) : s(seed) {}
  double uniform() {
    unsigned long long z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  }
};

}  // namespace

bool outerNormalsAxisExact(const mhy_config& o, std::string& err) {
  const long nx = o.nx, ny = o.ny;
  auto node = [&](long j, long i) { return P2{o.dx * (double)i, o.dy * (double)j}; };
  for (long j = 0; j < ny; ++j)
    for (long i = 0; i < nx; ++i) {
      const P2 bl = node(j, i), br = node(j, i + 1), tr = node(j + 1, i + 1), tl = node(j + 1, i);
      P2 xc = {0.0, 0.0};
      if (o.schema == MHY_SCHEMA_EUCCLHYD) {  // polygon centroid (eucclhyd_remap/Init.cc:106-145)
        const P2 X[4] = {bl, br, tr, tl};
        double r = 0.0;
        for (int q = 0; q < 4; ++q) r = r + cross(X[q], X[(q + 1) & 3]);
        const double vol = 0.5 * r;
        for (int q = 0; q < 4; ++q) xc = add(xc, mul(add(X[q], X[(q + 1) & 3]), cross(X[q], X[(q + 1) & 3])));
        xc = mul(xc, 1.0 / (6.0 * vol));
      } else {
        xc = add(xc, bl), xc = add(xc, br), xc = add(xc, tr), xc = add(xc, tl);
        xc = mul(xc, 0.25);
      }
      // bottom, left, right, top face centres and the axis component of
      // their outer normals
      const P2 fc[4] = {mul(add(bl, br), 0.5), mul(add(bl, tl), 0.5), mul(add(br, tr), 0.5), mul(add(tl, tr), 0.5)};
      const double want[4] = {-1.0, -1.0, 1.0, 1.0};
      for (int s = 0; s < 4; ++s) {
        const P2 d = sub(fc[s], xc);
        const double nrm = norm(d);
        const double comp = (s == 1 || s == 2) ? d.x / nrm : d.y / nrm;
        if (comp != want[s]) {
          err = "outer face normal of cell (" + std::to_string(j) + "," + std::to_string(i) +
                ") is not axis-exact; mesh not supported";
          return false;
        }
      }
    }
  return true;
}

int buildInitialState(const mhy_config& o, HostInit& s, std::string& err) {
  const long nx = o.nx, ny = o.ny;
  if (nx < 2 || ny < 2) {
    err = "meshes need at least 2 cells per direction";
    return MHY_ERR_ARG;
  }
  if (o.nbmat < 1 || o.nbmat > MHY_NBMATMAX) {
    err = "nbmat must be 1..3";
    return MHY_ERR_ARG;
  }
  const long C = nx * ny, N = (nx + 1) * (ny + 1);
  const int t = o.test_case;
  s.nx = nx, s.ny = ny, s.nbmat = o.nbmat;
  auto z = [&](std::vector<double>& v, long n) { v.assign(n, 0.0); };
  z(s.rho, C), z(s.p, C), z(s.c, C), z(s.e, C), z(s.vol, C), z(s.M, C);
  for (int k = 0; k < MHY_NBMATMAX; ++k) {
    z(s.rho_k[k], C), z(s.p_k[k], C), z(s.c_k[k], C), z(s.e_k[k], C);
    z(s.f_k[k], C), z(s.mu_k[k], C), z(s.M_k[k], C);
  }
  s.mixte.assign(C, 0), s.pure.assign(C, 0);
  z(s.ux, N), z(s.uy, N);

  // boundary conditions (Init.cc:9-95); the initialisers without a
  // reference counterpart use symmetry on all sides
  Boundary& b = s.bc;
  const int sym = MHY_BC_SYMMETRY, imp = MHY_BC_IMPOSED_VELOCITY;
  if (t == MHY_CASE_SOD_X || t == MHY_CASE_BI_SOD_X || t == MHY_CASE_SOD_Y || t == MHY_CASE_BI_SOD_Y ||
      t >= MHY_CASE_RIEMANN_PERTURBED) {
    b.left = b.right = b.top = b.bottom = sym;
  }
  if (t == MHY_CASE_ADVECTION_X || t == MHY_CASE_BI_ADVECTION_X || t == MHY_CASE_ADVECTION_VIT_X ||
      t == MHY_CASE_BI_ADVECTION_VIT_X) {
    b.right = imp, b.vright[0] = 1.0, b.vright[1] = 0.0, b.periodic_right = 1;
    b.left = imp, b.vleft[0] = 1.0, b.vleft[1] = 0.0;
    b.top = b.bottom = sym;
  }
  if (t == MHY_CASE_ADVECTION_Y || t == MHY_CASE_BI_ADVECTION_Y) {
    b.bottom = imp, b.vbottom[0] = 0.0, b.vbottom[1] = 1.0;
    b.top = imp, b.vtop[0] = 0.0, b.vtop[1] = 1.0;
    b.left = b.right = sym;
  }
  if (t == MHY_CASE_UNIT_TEST || t == MHY_CASE_BI_UNIT_TEST) {
    b.right = imp, b.vright[0] = 1.0, b.vright[1] = 0.0;
    b.left = imp, b.vleft[0] = 1.0, b.vleft[1] = 0.0;
    b.flux_left = 1;
    b.bottom = imp, b.vbottom[0] = 1.0, b.vbottom[1] = 0.0;
    b.top = imp, b.vtop[0] = 1.0, b.vtop[1] = 0.0;
  }

  auto node = [&](long j, long i) { return P2{o.dx * (double)i, o.dy * (double)j}; };
  auto cellCentre = [&](long j, long i) {
    P2 r = {0.0, 0.0};
    r = add(r, node(j, i)), r = add(r, node(j, i + 1)), r = add(r, node(j + 1, i + 1)), r = add(r, node(j + 1, i));
    return mul(r, 0.25);
  };
  const double gamma = o.gamma;
  auto setVelAll = [&](double vx, double vy) {
    for (long p = 0; p < N; ++p) s.ux[p] = vx, s.uy[p] = vy;
  };
  auto& f = s.f_k;
  auto& mu = s.mu_k;
  if (t == MHY_CASE_SOD_X || t == MHY_CASE_SOD_Y) {
    for (long c = 0; c < C; ++c) {
      const P2 xc = cellCentre(c / nx, c % nx);
      const double r = (t == MHY_CASE_SOD_X) ? xc.x : xc.y;
      f[0][c] = 1.0, mu[0][c] = 1.0;
      const double rho = (r < 0.5) ? 1.0 : 0.125, p = (r < 0.5) ? 1.0 : 0.1;
      s.rho[c] = rho, s.p[c] = p, s.rho_k[0][c] = rho, s.p_k[0][c] = p;
      // the reference's inverted sound speed sqrt(gamma*rho/p) (Init.cc:133-135)
      s.c_k[0][c] = std::sqrt(gamma * s.rho_k[0][c] / s.p_k[0][c]);
      s.c[c] = s.c_k[0][c];
    }
    setVelAll(0.0, 0.0);
  } else if (t == MHY_CASE_BI_SOD_X || t == MHY_CASE_BI_SOD_Y) {
    for (long c = 0; c < C; ++c) {
      const P2 xc = cellCentre(c / nx, c % nx);
      const double r = (t == MHY_CASE_BI_SOD_X) ? xc.x : xc.y;
      if (r < 0.5) {
        f[0][c] = 1.0, f[1][c] = 0.0, mu[0][c] = 1.0, mu[1][c] = 0.0;
        s.rho[c] = 1.0, s.p[c] = 1.0, s.rho_k[0][c] = 1.0, s.p_k[0][c] = 1.0;
      } else {
        f[0][c] = 0.0, f[1][c] = 1.0, mu[0][c] = 0.0, mu[1][c] = 1.0;
        s.rho[c] = 0.125, s.p[c] = 0.1, s.rho_k[1][c] = 0.125, s.p_k[1][c] = 0.1;
      }
      s.c_k[0][c] = std::sqrt(gamma * s.rho_k[0][c] / s.p_k[0][c]);
      s.c_k[1][c] = std::sqrt(gamma * s.rho_k[1][c] / s.p_k[1][c]);
      s.c[c] = smin(s.c_k[0][c], s.c_k[1][c]);
    }
    setVelAll(0.0, 0.0);
  } else if (t == MHY_CASE_ADVECTION_X || t == MHY_CASE_ADVECTION_Y || t == MHY_CASE_BI_ADVECTION_X ||
             t == MHY_CASE_BI_ADVECTION_Y || t == MHY_CASE_BI_ADVECTION_VIT_X || t == MHY_CASE_BI_ADVECTION_VIT_Y) {
    const bool vit = (t == MHY_CASE_BI_ADVECTION_VIT_X || t == MHY_CASE_BI_ADVECTION_VIT_Y);
    const bool bi = (t == MHY_CASE_BI_ADVECTION_X || t == MHY_CASE_BI_ADVECTION_Y) || vit;
    const bool isx = (t == MHY_CASE_ADVECTION_X || t == MHY_CASE_BI_ADVECTION_X || t == MHY_CASE_BI_ADVECTION_VIT_X);
    const double outer = vit ? 4.0 : 1.0;
    for (long c = 0; c < C; ++c) {
      const P2 xc = cellCentre(c / nx, c % nx);
      const double r = isx ? xc.x : xc.y;
      if (r < 0.3 || r > 0.5) {
        f[0][c] = 1.0, mu[0][c] = 1.0;
        s.rho[c] = outer, s.p[c] = 0.0, s.rho_k[0][c] = outer, s.p_k[0][c] = 0.0;
      } else if ((r > 0.3) && (r < 0.5)) {
        const int k = bi ? 1 : 0;
        f[k][c] = 1.0, mu[k][c] = 1.0;
        s.rho[c] = vit ? 4.0 : (bi ? 1.0 : 10.0);
        s.p[c] = 0.0;
        s.rho_k[k][c] = bi ? 1.0 : 10.0;
        s.p_k[k][c] = 0.0;
      }
      s.c_k[0][c] = 1.0, s.c_k[1][c] = 1.0;
      s.c[c] = smin(s.c_k[0][c], s.c_k[1][c]);
      s.e_k[0][c] = 1.0, s.e_k[1][c] = 1.0, s.e[c] = 1.0;
    }
    for (long j = 0; j <= ny; ++j)
      for (long i = 0; i <= nx; ++i) {
        const long p = j * (nx + 1) + i;
        double vx = isx ? 1.0 : 0.0, vy = isx ? 0.0 : 1.0;
        if (vit) {
          const double r = isx ? node(j, i).x : node(j, i).y;
          double sc = 1.0;
          if (r <= 0.3) sc = 0.2;
          else if ((r > 0.3) && (r <= 0.5)) sc = 4. * r - 1.;
          vx = vx * sc, vy = vy * sc;
        }
        s.ux[p] = vx, s.uy[p] = vy;
      }
  } else if (t == MHY_CASE_UNIT_TEST || t == MHY_CASE_BI_UNIT_TEST) {
    for (long c = 0; c < C; ++c) {
      f[0][c] = 1.0, f[1][c] = 0.0, mu[0][c] = 1.0, mu[1][c] = 0.0;
      s.p[c] = 1.0, s.p_k[0][c] = 1.0, s.p_k[1][c] = 0.0;
      s.c[c] = std::sqrt(gamma), s.c_k[0][c] = std::sqrt(gamma), s.c_k[1][c] = std::sqrt(gamma);
      s.rho_k[1][c] = 0.0;
      const double r = cellCentre(c / nx, c % nx).x;
      if (r < 0.4) s.rho[c] = 1.0, s.rho_k[0][c] = 1.0;
      else if (r > 0.4) s.rho[c] = 0.1, s.rho_k[0][c] = 0.1;
      if (r > 0.6) s.rho[c] = 0.1, s.rho_k[0][c] = 0.1;
    }
    setVelAll(1.0, 0.0);
  } else if (t >= MHY_CASE_RIEMANN_PERTURBED && t <= MHY_CASE_RIEMANN) {
    // per-material states, cut cells by sub-sampling, physical sound speed
    const int nm = o.nbmat;
    Rng rng(o.seed);
    const int nmodes = (int)o.param[2];
    std::vector<double> amp(nmodes > 0 ? nmodes : 1), ph(nmodes > 0 ? nmodes : 1);
    double asum = 0.0;
    if (t == MHY_CASE_RIEMANN_PERTURBED)
      for (int k = 0; k < nmodes; ++k) {
        amp[k] = 0.5 + 0.5 * rng.uniform();
        ph[k] = 2.0 * 3.14159265358979323846 * rng.uniform();
        asum += amp[k];
      }
    for (long c = 0; c < C; ++c) {
      const long j = c / nx, i = c - j * nx;
      double fr[MHY_NBMATMAX] = {0.0, 0.0, 0.0}, rho[MHY_NBMATMAX] = {0.0, 0.0, 0.0};
      double pr[MHY_NBMATMAX] = {0.0, 0.0, 0.0};
      if (t == MHY_CASE_RIEMANN_PERTURBED || t == MHY_CASE_RIEMANN) {
        const double x0 = o.param[0];
        if (t == MHY_CASE_RIEMANN_PERTURBED) {
          rho[0] = o.param[4], pr[0] = o.param[5], rho[1] = o.param[6], pr[1] = o.param[7];
        } else {
          rho[0] = o.param[1], pr[0] = o.param[2], rho[1] = o.param[3], pr[1] = o.param[4];
        }
        const double xl = node(j, i).x;
        const int ns = (t == MHY_CASE_RIEMANN_PERTURBED) ? 32 : 1;
        double cover = 0.0;
        for (int q = 0; q < ns; ++q) {
          const double ys = node(j, i).y + o.dy * ((double)q + 0.5) / (double)ns;
          double xi = x0;
          if (t == MHY_CASE_RIEMANN_PERTURBED) {
            double acc = 0.0;
            for (int k = 0; k < nmodes; ++k)
              acc += amp[k] * std::sin(2.0 * 3.14159265358979323846 * (double)(k + 1) * ys + ph[k]);
            xi = x0 + o.param[1] * acc / asum;
          }
          const double w = (xi - xl) / o.dx;
          cover += (w < 0.0) ? 0.0 : (w > 1.0 ? 1.0 : w);
        }
        fr[0] = cover / (double)ns;
        fr[1] = 1.0 - fr[0];
      } else if (t == MHY_CASE_POINT_EXPLOSION) {
        fr[0] = 1.0, rho[0] = o.param[0], pr[0] = o.param[1];
        if (j < 2 && i < 2) {
          const double vol = o.dx * o.dy;
          const double e = o.param[2] / 4.0 / (rho[0] * vol);
          pr[0] = (o.gammap[0] - 1.0) * rho[0] * e;
        }
      } else {  // MHY_CASE_TRIPLE_LAYOUT
        const P2 xc = cellCentre(j, i);
        const int k = (xc.x < 1.0) ? 0 : (xc.y >= 0.5 ? 1 : 2);
        const double r0[3] = {1.0, 0.125, 1.0}, p0[3] = {1.0, 0.1, 0.1};
        fr[k] = 1.0;
        rho[k] = r0[k] * (1.0 + o.param[0] * (2.0 * rng.uniform() - 1.0));
        pr[k] = p0[k];
      }
      double rmix = 0.0, pmix = 0.0;
      for (int k = 0; k < nm; ++k) {
        if (fr[k] <= 0.0) rho[k] = 0.0, pr[k] = 0.0;
        rmix += fr[k] * rho[k];
        pmix += fr[k] * pr[k];
      }
      double cmax = 0.0;
      for (int k = 0; k < nm; ++k) {
        f[k][c] = fr[k];
        mu[k][c] = fr[k] * rho[k] / rmix;
        s.rho_k[k][c] = rho[k];
        s.p_k[k][c] = pr[k];
        const double ck = (fr[k] > 0.0) ? std::sqrt(o.gammap[k] * pr[k] / rho[k]) : 0.0;
        s.c_k[k][c] = ck;
        cmax = smax(cmax, ck);
      }
      s.rho[c] = rmix, s.p[c] = pmix, s.c[c] = cmax;
    }
    setVelAll(0.0, 0.0);
  } else {
    err = "unknown test case " + std::to_string(t);
    return MHY_ERR_UNKNOWN_CASE;
  }

  for (long c = 0; c < C; ++c) {
    int matcell = 0, imatpure = -1;
    for (int k = 0; k < MHY_NBMATMAX; ++k)
      if (f[k][c] > o.threshold) matcell++, imatpure = k;
    s.mixte[c] = (matcell > 1) ? 1 : 0;
    s.pure[c] = (matcell > 1) ? -1 : imatpure;
  }
  // initSubVol (Init.cc:400-429): sum of the four corner sub-volumes
  for (long j = 0; j < ny; ++j)
    for (long i = 0; i < nx; ++i) {
      const P2 X[4] = {node(j, i), node(j, i + 1), node(j + 1, i + 1), node(j + 1, i)};
      const P2 xc = cellCentre(j, i);
      double v = 0.0;
      for (int q = 0; q < 4; ++q) {
        const P2 x2 = mul(add(X[(q + 3) & 3], X[q]), 0.5);
        const P2 x4 = mul(add(X[(q + 1) & 3], X[q]), 0.5);
        v = v + 0.5 * (cross(xc, x2) + cross(x2, X[q]) + cross(X[q], x4) + cross(x4, xc));
      }
      s.vol[j * nx + i] = v;
    }
  // computeCellMass (PhaseLagrange.cc:14-24)
  for (long c = 0; c < C; ++c) {
    s.M[c] = o.dx * o.dy * s.rho[c];
    for (int k = 0; k < o.nbmat; ++k) s.M_k[k][c] = mu[k][c] * s.M[c];
  }
  s.deltat_init = o.deltat_init;
  if (o.sans_lagrange == 0) {
    // initDeltaT (Init.cc:435-457): NaN sound speeds are skipped by the min
    double red = std::numeric_limits<double>::max();
    for (long c = 0; c < C; ++c) red = smin(red, 0.1 * std::sqrt(s.vol[c]) / s.c[c]);
    s.deltat_init = red * 1.0E-6;
    // initInternalEnergy (Init.cc:463-472)
    for (long c = 0; c < C; ++c) {
      for (int k = 0; k < MHY_NBMATMAX; ++k) s.e_k[k][c] = s.p_k[k][c] / ((o.gammap[k] - 1.0) * s.rho_k[k][c]);
      if (t >= MHY_CASE_RIEMANN_PERTURBED) {
        double e = 0.0;
        for (int k = 0; k < o.nbmat; ++k)
          if (f[k][c] > 0.0) e += mu[k][c] * s.e_k[k][c];
        s.e[c] = e;
      } else {
        s.e[c] = s.p[c] / ((o.gamma - 1.0) * s.rho[c]);
      }
    }
  }
  return MHY_OK;
}

}  // namespace mhy
