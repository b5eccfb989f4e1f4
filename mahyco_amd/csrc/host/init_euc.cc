// Initial state of an Eucclhyd problem (eucclhyd_remap/Init.cc:15-781):
// boundary conditions, the cell geometry of the set-up mesh (polygon
// centroid, volume, perimeter), the test-case initialisers and the initial
// time step, computed once on the host.
#include <cmath>

#include "init.hpp"

namespace mhy {

namespace {

struct Q2 {
  double x, y;
};
inline Q2 qadd(Q2 a, Q2 b) { return {a.x + b.x, a.y + b.y}; }
inline Q2 qsub(Q2 a, Q2 b) { return {a.x - b.x, a.y - b.y}; }
inline Q2 qmul(Q2 a, double s) { return {a.x * s, a.y * s}; }
inline double qcross(Q2 a, Q2 b) { return a.x * b.y - a.y * b.x; }
// norm with the dot product accumulated from 0.0 (includes/Freefunctions.h:13-23)
inline double qnorm(Q2 a) {
  double r = 0.0;
  r = r + a.x * a.x;
  r = r + a.y * a.y;
  return std::sqrt(r);
}

bool eucCase(int t) {
  switch (t) {
    case MHY_CASE_UNIT_TEST:
    case MHY_CASE_BI_UNIT_TEST:
    case MHY_CASE_SOD_X:
    case MHY_CASE_SOD_Y:
    case MHY_CASE_BI_SOD_X:
    case MHY_CASE_BI_SOD_Y:
    case MHY_CASE_BI_SHOCK_BUBBLE:
    case MHY_CASE_SEDOV:
    case MHY_CASE_TRIPLE_POINT:
    case MHY_CASE_BI_TRIPLE_POINT:
      return true;
    default:
      return false;
  }
}

}  // namespace

int buildEucInitialState(const mhy_config& o, HostInit& s, EucInit& u, std::string& err) {
  const int t = o.test_case, m = o.nbmat;
  if (!eucCase(t)) {
    err = "test case " + std::to_string(t) +
          " has no complete Eucclhyd initialisation (supported: Sod X/Y, bi-material Sod X/Y, Sedov, "
          "triple point, bi-material triple point, shock bubble, unit tests)";
    return MHY_ERR_UNKNOWN_CASE;
  }
  const long nx = o.nx, ny = o.ny, C = nx * ny, N = (nx + 1) * (ny + 1);
  s.nx = nx, s.ny = ny, s.nbmat = m;
  // boundary conditions (Init.cc:15-100); a case without a rule keeps the
  // defaults of includes/ConditionsLimites.h (code 0: zero velocity)
  Boundary& b = s.bc;
  const int sym = MHY_BC_SYMMETRY, imp = MHY_BC_IMPOSED_VELOCITY;
  auto set = [](double* d, double x, double y) { d[0] = x, d[1] = y; };
  if (t == MHY_CASE_SOD_X || t == MHY_CASE_SOD_Y || t == MHY_CASE_BI_SOD_X || t == MHY_CASE_BI_SOD_Y ||
      t == MHY_CASE_TRIPLE_POINT || t == MHY_CASE_BI_TRIPLE_POINT) {
    b.left = sym, b.right = sym, b.top = sym, b.bottom = sym;
    set(b.vleft, 0.0, 1.0), set(b.vright, 0.0, 1.0), set(b.vtop, 1.0, 0.0), set(b.vbottom, 1.0, 0.0);
  } else if (t == MHY_CASE_BI_SHOCK_BUBBLE) {
    b.left = sym, b.right = imp, b.top = sym, b.bottom = sym;
    set(b.vleft, 0.0, 1.0), set(b.vright, -124.824, 0.0), set(b.vtop, 1.0, 0.0), set(b.vbottom, 1.0, 0.0);
    b.flux_right = 1;
  } else if (t == MHY_CASE_SEDOV) {
    b.left = sym, b.right = imp, b.top = imp, b.bottom = sym;
    set(b.vleft, 0.0, 1.0), set(b.vright, 0.0, 0.0), set(b.vtop, 0.0, 0.0), set(b.vbottom, 1.0, 0.0);
  }
  // set-up mesh (CartesianMesh2DGenerator.cc:52-55) and cell geometry
  // (initMeshGeometryForCells, Init.cc:106-167)
  auto node = [&](long j, long i) { return Q2{o.dx * (double)i, o.dy * (double)j}; };
  u.xc.assign(C, 0.0), u.yc.assign(C, 0.0), u.ve.assign(C, 0.0), u.perim.assign(C, 0.0);
  u.vx.assign(C, 0.0), u.vy.assign(C, 0.0);
  s.rho.assign(C, 0.0), s.e.assign(C, 0.0);
  for (int k = 0; k < m; ++k) {
    s.rho_k[k].assign(C, 0.0), s.e_k[k].assign(C, 0.0), s.f_k[k].assign(C, 0.0), s.mu_k[k].assign(C, 0.0);
  }
  s.mixte.assign(C, 0), s.pure.assign(C, 0);
  s.ux.assign(N, 0.0), s.uy.assign(N, 0.0);
  const double* g = o.gammap;
  const Q2 Xb = {0.320, 0.04};
  const double rb = 0.025;
  for (long j = 0; j < ny; ++j)
    for (long i = 0; i < nx; ++i) {
      const long c = j * nx + i;
      const Q2 X[4] = {node(j, i), node(j, i + 1), node(j + 1, i + 1), node(j + 1, i)};
      double r11 = 0.0;
      for (int q = 0; q < 4; ++q) r11 = r11 + qcross(X[q], X[(q + 1) & 3]);
      const double vol = 0.5 * r11;
      Q2 r12 = {0.0, 0.0};
      for (int q = 0; q < 4; ++q) r12 = qadd(r12, qmul(qadd(X[q], X[(q + 1) & 3]), qcross(X[q], X[(q + 1) & 3])));
      const Q2 xc = qmul(r12, 1.0 / (6.0 * vol));
      double r13 = 0.0;
      for (int q = 0; q < 4; ++q) r13 = r13 + qnorm(qsub(X[q], X[(q + 1) & 3]));
      u.xc[c] = xc.x, u.yc[c] = xc.y, u.ve[c] = vol, u.perim[c] = r13;

      // internal energy (Init.cc:194-385), three material slots as in the
      // reference, the unused ones discarded below
      double e = 0.0, ek[3] = {0.0, 0.0, 0.0};
      if (t == MHY_CASE_UNIT_TEST || t == MHY_CASE_BI_UNIT_TEST) {
        const double p0 = 1., rho0 = 1., gamma = 1.4;
        const double e0 = p0 / ((gamma - 1.0) * rho0);
        e = e0, ek[0] = ek[1] = ek[2] = e0;
      } else if (t == MHY_CASE_SOD_X || t == MHY_CASE_SOD_Y) {
        const double r = (t == MHY_CASE_SOD_X) ? xc.x : xc.y;
        const double pI = (r <= 0.5) ? 1.0 : 0.1, rI = (r <= 0.5) ? 1.0 : 0.125;
        e = ek[0] = pI / ((g[0] - 1.0) * rI);
      } else if (t == MHY_CASE_BI_SOD_X || t == MHY_CASE_BI_SOD_Y) {
        const double r = (t == MHY_CASE_BI_SOD_X) ? xc.x : xc.y;
        if (r <= 0.5) e = ek[0] = 1.0 / ((g[0] - 1.0) * 1.0);
        else e = ek[1] = 0.1 / ((g[1] - 1.0) * 0.125);
      } else if (t == MHY_CASE_BI_SHOCK_BUBBLE) {
        const double r = std::sqrt((xc.x - Xb.x) * (xc.x - Xb.x) + (xc.y - Xb.y) * (xc.y - Xb.y));
        if (r <= rb) e = ek[1] = 1.e5 / ((g[0] - 1.0) * 0.182);
        else e = ek[0] = 1.e5 / ((g[0] - 1.0) * 1.0);
      } else if (t == MHY_CASE_SEDOV) {
        bool center = false;
        for (int q = 0; q < 4; ++q)
          if (qnorm(X[q]) < o.threshold) center = true;
        const double e1 = 1.e-6 / ((g[0] - 1.0) * 1.);
        e = center ? e1 + 0.244816 / (o.dx * o.dy) : e1;
        ek[0] = e;
      } else if (t == MHY_CASE_TRIPLE_POINT) {
        double pI, rI;
        if (xc.x <= 0.01) pI = 1.0, rI = 1.0;
        else if (xc.y <= 0.015) pI = 0.1, rI = 1.;
        else pI = 0.1, rI = 0.1;
        e = ek[0] = pI / ((g[0] - 1.0) * rI);
      } else {  // bi-material triple point
        if (xc.x <= 0.01) e = ek[0] = 1.0 / ((g[0] - 1.0) * 1.0);
        else if (xc.y <= 0.015) e = ek[1] = 0.1 / ((g[1] - 1.0) * 1.);
        else e = ek[2] = 0.1 / ((g[2] - 1.0) * 0.1);
      }
      // cell velocity (Init.cc:391-423)
      if (t == MHY_CASE_BI_SHOCK_BUBBLE && xc.x >= 0.60) u.vx[c] = -124.824;
      if (t == MHY_CASE_UNIT_TEST || t == MHY_CASE_BI_UNIT_TEST) u.vx[c] = 1.0;
      // density and fractions (Init.cc:429-671)
      double fr[3] = {0.0, 0.0, 0.0}, mf[3] = {0.0, 0.0, 0.0}, rk[3] = {0.0, 0.0, 0.0}, rho = 0.0;
      auto only = [&](int k, double r, double rkv) {
        fr[k] = 1., mf[k] = 1., rho = r, rk[k] = rkv;
      };
      if (t == MHY_CASE_UNIT_TEST) {
        only(0, 1.0, 1.0);
      } else if (t == MHY_CASE_BI_UNIT_TEST) {
        if (xc.x <= 0.5) only(0, 1.0, 1.0);
        else only(1, 1., 1.0);
      } else if (t == MHY_CASE_SOD_X || t == MHY_CASE_SOD_Y) {
        const double r = (t == MHY_CASE_SOD_X) ? xc.x : xc.y;
        const double d = (r <= 0.5) ? 1.0 : 0.125;
        only(0, d, d);
        rk[1] = d;  // the reference sets the second slot too (unused)
      } else if (t == MHY_CASE_BI_SOD_X || t == MHY_CASE_BI_SOD_Y) {
        const double r = (t == MHY_CASE_BI_SOD_X) ? xc.x : xc.y;
        if (r <= 0.5) only(0, 1.0, 1.0);
        else only(1, 0.125, 0.125);
      } else if (t == MHY_CASE_BI_SHOCK_BUBBLE) {
        const double r = std::sqrt((xc.x - Xb.x) * (xc.x - Xb.x) + (xc.y - Xb.y) * (xc.y - Xb.y));
        if (r <= rb) only(1, 0.182, 0.182);
        else only(0, 1.0, 1.0);
      } else if (t == MHY_CASE_TRIPLE_POINT) {
        // the material density stays unset (zero), as in the reference
        only(0, (xc.x <= 0.01) ? 1.0 : ((xc.y <= 0.015) ? 1.0 : 0.1), 0.0);
      } else if (t == MHY_CASE_BI_TRIPLE_POINT) {
        if (xc.x <= 0.01) only(0, 1.0, 1.0);
        else if (xc.y <= 0.015) only(1, 1.0, 1.0);
        else only(2, 0.1, 0.1);
      } else {  // Sedov
        rho = 1.0, rk[0] = 1.0;
      }
      if (m == 1) fr[0] = 1., fr[1] = fr[2] = 0., mf[0] = 1., mf[1] = mf[2] = 0.;
      s.rho[c] = rho, s.e[c] = e;
      for (int k = 0; k < m; ++k) {
        s.rho_k[k][c] = rk[k], s.e_k[k][c] = ek[k], s.f_k[k][c] = fr[k], s.mu_k[k][c] = mf[k];
      }
      int matcell = 0, imatpure = -1;
      for (int k = 0; k < 3; ++k)
        if (fr[k] > o.threshold) matcell++, imatpure = k;
      s.mixte[c] = (matcell > 1) ? 1 : 0;
      s.pure[c] = (matcell > 1) ? -1 : imatpure;
    }
  u.outvx = u.vx, u.outvy = u.vy;
  // initial time step of the case (setUpTimeLoopN, Init.cc:721-760); the
  // time state is value-initialised, so a case without one starts at 0
  s.deltat_init = 0.0;
  if (t == MHY_CASE_SOD_X || t == MHY_CASE_SOD_Y || t == MHY_CASE_BI_SOD_X || t == MHY_CASE_BI_SOD_Y ||
      t == MHY_CASE_SEDOV)
    s.deltat_init = 1.0e-4;
  else if (t == MHY_CASE_BI_SHOCK_BUBBLE)
    s.deltat_init = 1.0e-7;
  else if (t == MHY_CASE_TRIPLE_POINT || t == MHY_CASE_BI_TRIPLE_POINT)
    s.deltat_init = 1.0e-5;
  return MHY_OK;
}

}  // namespace mhy
