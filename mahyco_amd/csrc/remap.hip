// Directionally split remap on CDNA4: one sweep kernel per direction
// (face gradients, limited cell gradients, PenteBorne swept-volume fluxes or
// upwind face values, conservative update), the nodal (dual) momentum
// remap, and the final remap with the EOS.
//
// Layout: structure-of-arrays FP64 in HBM, one thread per cell (node) with
// the 5-cell stencil along the sweep read through the L1/L2 caches; the
// compact remap vector holds the 3m+1 live components (volume, mass and
// energy per material, pseudo-viscosity content).
#include "common.hpp"
#include "kernels.hpp"

namespace mhy {

namespace {

__device__ __forceinline__ void raise(Scalars* sc, int code) { atomicCAS(&sc->err, 0, code); }

// remap/UtilesRemap.cc:9-23
__device__ __forceinline__ double fluxLimiter(int lim, double r) {
  if (lim == MHY_LIM_MINMOD) return smax(0.0, smin(1.0, r));
  if (lim == MHY_LIM_SUPERBEE) return smax(0.0, smax(smin(2.0 * r, 1.0), smin(r, 2.0)));
  if (lim == MHY_LIM_VANLEER) return (r <= 0.0) ? 0.0 : 2.0 * r / (1.0 + r);
  return 0.0;
}

// remap/UtilesRemap.cc:25-72
__device__ double fluxLimiterPP(int lim, double gradplus, double gradmoins, double yplus, double ymoins, double h0,
                                double hplus, double hmoins) {
  double grady = 0.0;
  if (gradplus * gradmoins < 0.0) return 0.;
  if (lim == MHY_LIM_MINMODG) {
    grady = ((yplus - ymoins) > 0.) ? smin(fabs(gradplus), fabs(gradmoins)) : -smin(fabs(gradplus), fabs(gradmoins));
  } else if (lim == MHY_LIM_SUPERBEEG) {
    grady = ((yplus - ymoins) > 0.) ? smax(fabs(gradplus), fabs(gradmoins)) : -smax(fabs(gradplus), fabs(gradmoins));
  } else if (lim == MHY_LIM_VANLEERG) {
    double lp = (h0 / 2. + hplus) / (h0 + hplus + hmoins);
    double lm = (h0 / 2. + hmoins) / (h0 + hplus + hmoins);
    if ((lp * gradplus + lm * gradmoins) != 0.) grady = gradplus * gradmoins / (lp * gradplus + lm * gradmoins);
    else grady = 0.;
  } else if (lim == MHY_LIM_ULTRABEEG) {
    grady = (yplus - ymoins) / h0;
  } else if (lim == MHY_LIM_ARITHMETICG) {
    double lp = (h0 / 2. + hplus) / (h0 + hplus + hmoins);
    double lm = (h0 / 2. + hmoins) / (h0 + hplus + hmoins);
    grady = lm * gradplus + lp * gradmoins;
  }
  double gMp = gradplus * (h0 + hplus) / h0;
  double gMm = gradmoins * (h0 + hmoins) / h0;
  double gM = smin(fabs(gMp), fabs(gMm));
  if ((yplus - ymoins) > 0.) grady = smin(fabs(gM), fabs(grady));
  else grady = -smin(fabs(gM), fabs(grady));
  return grady;
}

// remap/UtilesRemap-Impl.h:99-126, one component
__device__ __forceinline__ double limitedGradient(int lim, double gp, double gm, double phi, double phip,
                                                  double phim, double h0, double hp, double hm) {
  (void)phi;
  if (lim < MHY_LIM_MINMODG)
    return (fluxLimiter(lim, divideNoExcept(gp, gm)) * gm + fluxLimiter(lim, divideNoExcept(gm, gp)) * gp) / 2.0;
  return fluxLimiterPP(lim, gp, gm, phip, phim, h0, hp, hm);
}

// remap/UtilesRemap.cc:74-125 (both thresholds at once)
__device__ void computeY0(int lim, double yplus, double ymoins, double h0, double hplus, double hmoins, double& y0plus,
                          double& y0moins) {
  y0plus = 0., y0moins = 0.;
  if (lim == MHY_LIM_MINMODG || lim == MHY_LIM_MINMOD) {
    y0plus = yplus, y0moins = ymoins;
  } else if (lim == MHY_LIM_SUPERBEEG || lim == MHY_LIM_SUPERBEE) {
    y0plus = ((h0 + hmoins) * yplus + h0 * ymoins) / (2 * h0 + hmoins);
    y0moins = ((h0 + hplus) * ymoins + h0 * yplus) / (2 * h0 + hplus);
  } else if (lim == MHY_LIM_VANLEERG || lim == MHY_LIM_VANLEER) {
    double a = smin(yplus, ymoins), b = smax(yplus, ymoins);
    double xplus = (h0 * h0 + 3 * h0 * hmoins + 2 * hmoins * hmoins) * yplus;
    double xmoins = (h0 * h0 + 3 * h0 * hplus + 2 * hplus * hplus) * ymoins;
    xplus += (h0 * h0 - h0 * hplus - 2 * hplus * hplus + 2 * h0 * hmoins) * ymoins;
    xmoins += (h0 * h0 - h0 * hmoins - 2 * hmoins * hmoins + 2 * h0 * hplus) * yplus;
    xplus /= (2 * h0 * h0 + 5 * h0 * hmoins + 2 * hmoins * hmoins - h0 * hplus - 2 * hplus * hplus);
    xmoins /= (2 * h0 * h0 + 5 * h0 * hplus + 2 * hplus * hplus - h0 * hmoins - 2 * hmoins * hmoins);
    y0plus = smin(smax(xplus, a), b);
    y0moins = smin(smax(xmoins, a), b);
  } else if (lim == MHY_LIM_ULTRABEEG) {
    y0plus = (yplus + ymoins) / 2.;
    y0moins = (yplus + ymoins) / 2.;
  } else if (lim == MHY_LIM_ARITHMETICG) {
    y0plus = ((h0 + hmoins + hplus) * yplus + h0 * ymoins) / (2 * h0 + hmoins + hplus);
    y0moins = ((h0 + hmoins + hplus) * ymoins + h0 * yplus) / (2 * h0 + hmoins + hplus);
  } else if (lim == 3000) {
    y0plus = yplus, y0moins = ymoins;
  }
}

// remap/UtilesRemap.cc:177-188
__device__ __forceinline__ double INT2Y(double X, double x0, double y0, double x1, double y1) {
  double flux = 0.;
  if (fabs(x1 - x0) > 1.e-14) {
    double eta = smin(smax(0., (X - x0) / (x1 - x0)), 1.);
    flux = (y0 + 0.5 * eta * (y1 - y0)) * (x1 - x0) * eta;
  }
  return flux;
}

// Sub-cell profile of one component (UtilesRemap-Impl.h:164-180 with
// UtilesRemap.cc:127-166): breakpoints xg, xd and values yg, yd.
struct Profile {
  double xg, xd, yg, yd;
};
__device__ Profile profile(int lim, double thr, double y0, double yplus, double ymoins, double h0, double hplus,
                           double hmoins, double grady) {
  double y0p, y0m;
  computeY0(lim, yplus, ymoins, h0, hplus, hmoins, y0p, y0m);
  Profile pr;
  double xplus = 1.;
  if (fabs(y0p - yplus) > thr) xplus = (y0 - yplus) / (y0p - yplus) - 1. / 2.;
  double xmoins = 1.;
  if (fabs(y0m - ymoins) > thr) xmoins = (y0 - ymoins) / (y0m - ymoins) - 1. / 2.;
  pr.xd = +h0 * smin(smax(xplus, -1. / 2.), 1. / 2.);
  pr.xg = -h0 * smin(smax(xmoins, -1. / 2.), 1. / 2.);
  const double xtd = y0 + h0 / 2 * grady;
  const double xtg = y0 - h0 / 2 * grady;
  const double ad = smin(yplus, 2. * y0m - ymoins), bd = smax(yplus, 2. * y0m - ymoins);
  const double ag = smin(ymoins, 2. * y0p - yplus), bg = smax(ymoins, 2. * y0p - yplus);
  pr.yd = smin(smax(xtd, ad), bd);
  pr.yg = smin(smax(xtg, ag), bg);
  return pr;
}

// Swept integral of the three-segment profile between a and b
// (UtilesRemap-Impl.h:185-200 and :230-247).
__device__ __forceinline__ double sweptIntegral(const Profile& pr, double h0, double ymoins, double yplus, double a,
                                                double b) {
  // b is the upper, a the lower integration bound; the reference forms
  // (f(b)-f(a)) per segment and sums the three differences
  const double f1b = INT2Y(b, -h0 / 2., ymoins, pr.xg, pr.yg), f1a = INT2Y(a, -h0 / 2., ymoins, pr.xg, pr.yg);
  const double f2b = INT2Y(b, pr.xg, pr.yg, pr.xd, pr.yd), f2a = INT2Y(a, pr.xg, pr.yg, pr.xd, pr.yd);
  const double f3b = INT2Y(b, pr.xd, pr.yd, h0 / 2., yplus), f3a = INT2Y(a, pr.xd, pr.yd, h0 / 2., yplus);
  return smax(((f1b - f1a) + (f2b - f2a) + (f3b - f3a)), 0.);
}

// Finishing of a PenteBorne flux vector (UtilesRemap-Impl.h:259-313 and
// :436-487): derived mass / energy / pseudo components.
__device__ void finishPP(const Opts& o, bool pureVariant, const double* phi, double* F) {
  const int m = o.nbmat;
  if (o.pp_complet == 1) {
    double sfv = 0.;
    for (int k = 0; k < m; ++k) sfv += F[k];
    if (fabs(sfv) > o.threshold) {
      // the summed mass flux only scales the dropped momentum components
      for (int k = 0; k < m; ++k) {
        if (!pureVariant) F[m + k] = (F[m + k] / sfv) * F[k];
        F[2 * m + k] = (F[2 * m + k] / sfv) * F[m + k];
      }
      F[3 * m] = phi[3 * m] * sfv;
    } else {
      for (int q = 0; q < o.nc; ++q) F[q] = 0.0;
    }
  } else {
    double sfv = 0.;
    for (int k = 0; k < m; ++k) {
      if (!pureVariant) F[m + k] = phi[m + k] * F[k];
      F[2 * m + k] = phi[2 * m + k] * F[m + k];
      sfv += F[k];
    }
    F[3 * m] = phi[3 * m] * sfv;
  }
}

// Stencil of 5 cells along the sweep around the current cell.
struct Stencil {
  long id[5];     // cell ids (clamped to the centre when outside)
  bool in[5];     // inside the domain
  double phi[5][NCMAX];
  double xc[5];   // along-sweep coordinate of the Lagrange centroid
  double H[5];    // Lagrange cell width across the sweep faces
  int mixte[5], pure[5];
};

__device__ __forceinline__ D2 lagrangeNode(const KP& kp, long j, long i) {
  const long p = j * (kp.o.nx + 1) + i;
  const D2 e = eulerNode(kp.o, j, i);
  const double dt1 = kp.f.sc->dt[kp.nxt];
  return {e.x + kp.f.ux[kp.nxt][p] * dt1, e.y + kp.f.uy[kp.nxt][p] * dt1};
}

// Phi of one cell at the start of the sweep: U/V (PrepareRemap.cc:171-173,
// PhaseRemap1.cc:548-550) or the PenteBorne form (PrepareRemap.cc:137-170,
// PhaseRemap1.cc:513-546; its pseudo component is the Lagrange-phase value).
__device__ void cellPhi(const KP& kp, double* const* U, long c, double* phi) {
  const int m = kp.o.nbmat, nc = kp.o.nc;
  if (kp.o.pp == 1) {
    double u[NCMAX];
    for (int q = 0; q < 3 * m; ++q) u[q] = U[q][c];
    double sv = 0.0;
    for (int k = 0; k < m; ++k) sv += u[k];
    for (int k = 0; k < m; ++k) {
      phi[k] = u[k] / sv;
      phi[m + k] = (u[k] != 0.) ? u[m + k] / u[k] : 0.0;
    }
    for (int k = 0; k < m; ++k) phi[2 * m + k] = (u[m + k] != 0.) ? u[2 * m + k] / u[m + k] : 0.0;
    phi[3 * m] = kp.f.qL[c];
  } else {
    const double V = kp.f.VL[c];
    for (int q = 0; q < nc; ++q) phi[q] = U[q][c] / V;
  }
}

// mixed / pure indicators from a Phi vector (PhaseRemap1.cc:552-566)
__device__ __forceinline__ void flagsOf(int m, const double* phi, int& mixte, int& pure) {
  int matcell = 0, imatpure = -1;
  for (int k = 0; k < m; ++k)
    if (phi[k] > 0.) matcell++, imatpure = k;
  if (matcell > 1) mixte = 1, pure = -1;
  else mixte = 0, pure = imatpure;
}

}  // namespace

// One remap sweep (Remap::computeGradPhiFace{1,2}, computeGradPhi{1,2},
// computeUpwindFaceQuantitiesForProjection{1,2}, computeUremap{1,2};
// remap/PhaseRemap1.cc, remap/PhaseRemap2.cc) for the cell of this thread.
// XD: sweep along x (vertical faces) or along y (horizontal faces).
template <bool XD>
__global__ __launch_bounds__(256) void k_sweep(KP kp) {
  const Opts& o = kp.o;
  const Fields& f = kp.f;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= kp.ncell) return;
  const int m = o.nbmat, nc = o.nc;
  const long nx = o.nx, ny = o.ny;
  const long j = c / nx, i = c - j * nx;
  const long a = XD ? i : j, na = XD ? nx : ny, stride = XD ? 1 : nx;
  double* const* Uin = (kp.sweep == 1) ? f.UL : f.U1;
  double* const* Uout = (kp.sweep == 1) ? f.U1 : f.U2;
  const double dtn = f.sc->dt[kp.cur];
  const bool pp = (o.pp == 1);
  const bool needH = pp || o.limiter >= MHY_LIM_MINMODG || (pp && o.limiter_pure >= MHY_LIM_MINMODG);

  Stencil st;
  for (int d = 0; d < 5; ++d) {
    const long b = a + d - 2;
    st.in[d] = (b >= 0 && b < na);
    st.id[d] = st.in[d] ? c + (d - 2) * stride : c;
  }
  for (int d = 0; d < 5; ++d) {
    if (!st.in[d]) continue;
    const long cd = st.id[d];
    cellPhi(kp, Uin, cd, st.phi[d]);
    st.xc[d] = XD ? f.XcL[cd] : f.YcL[cd];
    if (needH) {
      const long jd = cd / nx, id = cd - jd * nx;
      double h = 0.;
      if (XD) {  // bottom then top face lengths (PhaseRemap1.cc:60-70)
        h += 0.5 * d2norm(d2sub(lagrangeNode(kp, jd, id + 1), lagrangeNode(kp, jd, id)));
        h += 0.5 * d2norm(d2sub(lagrangeNode(kp, jd + 1, id + 1), lagrangeNode(kp, jd + 1, id)));
      } else {  // right then left face lengths (PhaseRemap1.cc:108-118)
        h += 0.5 * d2norm(d2sub(lagrangeNode(kp, jd + 1, id + 1), lagrangeNode(kp, jd, id + 1)));
        h += 0.5 * d2norm(d2sub(lagrangeNode(kp, jd + 1, id), lagrangeNode(kp, jd, id)));
      }
      st.H[d] = h;
    } else {
      st.H[d] = 0.0;
    }
    if (pp && o.pp_mixte == 1) {
      if (kp.sweep == 1) st.mixte[d] = f.mixte[cd], st.pure[d] = f.pure[cd];
      else flagsOf(m, st.phi[d], st.mixte[d], st.pure[d]);
    }
  }
  // clamp missing neighbours to the cell itself (PhaseRemap1.cc:157-158)
  auto S = [&](int d) { return st.in[d] ? d : 2; };

  // face normal velocity of face b (between along cells b-1 and b), from the
  // end-of-Lagrange node velocities (PrepareRemap.cc:50-62)
  auto fnv = [&](long b) {
    long n1, n2;
    if (XD) n1 = j * (nx + 1) + b, n2 = (j + 1) * (nx + 1) + b;
    else n1 = b * (nx + 1) + i, n2 = b * (nx + 1) + i + 1;
    const double* u = XD ? f.ux[kp.nxt] : f.uy[kp.nxt];
    return 0.0 + ((0.0 + u[n1]) + u[n2]) * 0.5;
  };
  // Euler length of face b (Init.cc:528-533)
  auto faceLen = [&](long b) {
    D2 e1, e2;
    if (XD) e1 = eulerNode(o, j, b), e2 = eulerNode(o, j + 1, b);
    else e1 = eulerNode(o, b, i), e2 = eulerNode(o, b, i + 1);
    return d2norm(d2sub(e2, e1));
  };
  // face gradient of component q at the face between stencil slots d-1, d
  auto gradFace = [&](int d, int q) -> double {
    if (!(st.in[d] && st.in[d - 1])) return 0.0;  // outer faces keep 0 (Q6)
    if (XD) return (st.phi[d][q] - st.phi[d - 1][q]) / (st.xc[d] - st.xc[d - 1]);
    return (st.phi[d - 1][q] - st.phi[d][q]) / (st.xc[d - 1] - st.xc[d]);
  };
  // limited gradient of component q of the cell in slot d (1..3)
  auto vpOf = [&](int d) {
    if (!(pp && o.pp_mixte == 1)) return false;
    const int dm = S(d - 1), dp = S(d + 1);
    return st.mixte[d] == 0 && st.mixte[dp] == 0 && st.mixte[dm] == 0 && st.pure[d] == st.pure[dp] &&
           st.pure[d] == st.pure[dm];
  };
  auto grad = [&](int d, int q, bool vp) -> double {
    if (o.order <= 1) return 0.0;
    const int lim = (pp && vp) ? o.limiter_pure : o.limiter;
    const double gMinus = gradFace(d, q), gPlus = gradFace(d + 1, q);
    const double gp = XD ? gPlus : gMinus, gm = XD ? gMinus : gPlus;
    const int dm = S(d - 1), dp = S(d + 1);
    return limitedGradient(lim, gp, gm, st.phi[d][q], st.phi[dp][q], st.phi[dm][q], st.H[d], st.H[dp], st.H[dm]);
  };

  // face values phi_f and flux scales of the minus / plus faces: the flux
  // of a face is coef * phi_f, entering the cell with the sign of its outer
  // normal (-1 on the minus face, +1 on the plus face)
  double phiM[NCMAX], phiP[NCMAX];
  const bool hasM = (a >= 1), hasP = (a + 1 < na);
  const double Lm = faceLen(a), Lp = faceLen(a + 1);
  const double vmAll = fnv(a), vpAll = fnv(a + 1);
  double coefM, coefP;
  if (!pp) {
    coefM = vmAll * Lm * dtn;  // (n.e * u_f) * L_f * dt with n.e = +-1
    coefP = vpAll * Lp * dtn;
    // upwind face values (UtilesRemap-Impl.h:491-501)
    auto faceValue = [&](long b, int dBack, int dFront, double v, double* phiF) {
      // back/front cells of the face: x: left/right; y: above/below
      const double dxl = st.xc[dFront] - st.xc[dBack];
      const int up = (v * dxl > 0.0) ? dBack : dFront;
      const double xf = XD ? 0.5 * (o.dx * (double)b + o.dx * (double)b) : 0.5 * (o.dy * (double)b + o.dy * (double)b);
      const double s = 0.0 + (xf - st.xc[up]);
      const bool vpu = vpOf(up);
      for (int q = 0; q < nc; ++q) phiF[q] = st.phi[up][q] + grad(up, q, vpu) * s;
    };
    if (hasM) {
      if (XD) faceValue(a, 1, 2, vmAll, phiM);
      else faceValue(a, 2, 1, vmAll, phiM);
    }
    if (hasP) {
      if (XD) faceValue(a + 1, 2, 3, vpAll, phiP);
      else faceValue(a + 1, 3, 2, vpAll, phiP);
    }
  } else {
    coefM = Lm;  // n.e * L_f
    coefP = Lp;
    // PenteBorne: phi_f = Av(minus cell) - Ar(plus cell) (PhaseRemap1.cc:361-364,
    // :420-423) from the swept-volume integrals of computeFluxPP
    const int lim = o.limiter;
    const int limPure = o.limiter_pure;
    const bool complet = (o.pp_complet == 1);
    const bool pb = (o.methode == 2);
    auto ppCell = [&](int d, double vminus, double vplus, double* Ar, double* Av, double* Dual) {
      const bool vp = vpOf(d);
      const int dm = S(d - 1), dp = S(d + 1);
      const int lz = vp ? limPure : lim;
      const double h0 = st.H[d], hp = st.H[dp], hm = st.H[dm];
      const double ar = -vminus;  // outer normal . e = -1 on the minus face
      const double av = vplus;
      const double dualv = 0.5 * (vminus + vplus);
      const double ppvR = 0.5 * (ar + fabs(ar)) * dtn;
      const double ppvV = 0.5 * (av + fabs(av)) * dtn;
      const double ppdv = 0.5 * (dualv + fabs(dualv)) * dtn;
      for (int q = 0; q < nc; ++q) Ar[q] = Av[q] = 0.0;
      if (Dual)
        for (int q = 0; q < nc; ++q) Dual[q] = 0.0;
      // components whose swept integral is used; the others are derived
      const int nInt = complet ? 3 * m : (vp ? 2 * m : m);
      const int qEnd = (Dual && nInt < 2 * m) ? 2 * m : nInt;
      for (int q = 0; q < qEnd; ++q) {
        const double phi = st.phi[d][q], phip = st.phi[dp][q], phim = st.phi[dm][q];
        const double g = grad(d, q, vp);
        const Profile pr = profile(lz, o.threshold, phi, phip, phim, h0, hp, hm, g);
        const bool flat = ((phip - phi) * (phim - phi)) >= 0.;
        if (q < nInt) {
          Ar[q] = flat ? phi * ppvR : sweptIntegral(pr, h0, phim, phip, -h0 / 2., -h0 / 2. + ppvR);
          Av[q] = flat ? phi * ppvV : sweptIntegral(pr, h0, phim, phip, h0 / 2. - ppvV, h0 / 2.);
        }
        if (Dual && q >= m && q < 2 * m) Dual[q] = flat ? phi * ppdv : sweptIntegral(pr, h0, phim, phip, 0., ppdv);
      }
      finishPP(o, vp, st.phi[d], Ar);
      finishPP(o, vp, st.phi[d], Av);
      if (ppvR == 0.)
        for (int q = 0; q < nc; ++q) Ar[q] = 0.0;
      if (ppvV == 0.)
        for (int q = 0; q < nc; ++q) Av[q] = 0.0;
      if (Dual && ppdv == 0.)
        for (int q = 0; q < nc; ++q) Dual[q] = 0.0;
    };
    if (o.order <= 1) {
      // the PenteBorne fluxes are only formed for order > 1; the face values
      // then stay at their zero initial value
      for (int q = 0; q < nc; ++q) phiM[q] = phiP[q] = 0.0;
      if (pb)
        for (int k = 0; k < m; ++k) f.DF[k][c] = 0.0;
    } else {
      double ArC[NCMAX], AvC[NCMAX], DualC[NCMAX];
      ppCell(2, vmAll, vpAll, ArC, AvC, pb ? DualC : nullptr);
      if (pb) {
        const double w = (Lm + Lp) * 0.5;
        for (int k = 0; k < m; ++k) f.DF[k][c] = DualC[m + k] * w;
      }
      if (hasM) {
        double ArM[NCMAX], AvM[NCMAX];
        ppCell(1, fnv(a - 1), vmAll, ArM, AvM, nullptr);
        for (int q = 0; q < nc; ++q) phiM[q] = AvM[q] - ArC[q];
      }
      if (hasP) {
        double ArP[NCMAX], AvP[NCMAX];
        ppCell(3, vpAll, fnv(a + 2), ArP, AvP, nullptr);
        for (int q = 0; q < nc; ++q) phiP[q] = AvC[q] - ArP[q];
      }
    }
  }
  double Gm[NCMAX], Gp[NCMAX];
  if (hasM)
    for (int q = 0; q < nc; ++q) Gm[q] = coefM * phiM[q];
  if (hasP)
    for (int q = 0; q < nc; ++q) Gp[q] = coefP * phiP[q];

  // computeUremap (PhaseRemap1.cc:468-503): neighbours below, above, left,
  // right; the faces across the sweep contribute 0
  double red[NCMAX];
  for (int q = 0; q < nc; ++q) red[q] = 0.0;
  if (hasM)
    for (int q = 0; q < nc; ++q) red[q] = red[q] + (-Gm[q]);
  if (hasP)
    for (int q = 0; q < nc; ++q) red[q] = red[q] + Gp[q];
  if (o.flux_any) {
    // remap/BoundaryFluxes.cc:14-120: flux through one face of the cell with
    // the face value of the opposite face; the first matching rule applies
    int rule = 0;  // 1: through the minus face, 2: through the plus face
    if (!XD && o.flux_bottom == 1 && c < nx) rule = 1;
    else if (!XD && o.flux_top == 1 && c <= nx * (ny - 1) && c < nx * ny) rule = 2;
    else if (XD && o.flux_left == 1 && i == 0) rule = 1;
    else if (XD && o.flux_right == 1 && i == nx - 1) rule = 2;
    if (rule != 0) {
      const bool viaMinus = (rule == 1);
      const double* phiOpp = viaMinus ? phiP : phiM;
      const bool oppOk = viaMinus ? hasP : hasM;
      const double vb = viaMinus ? vmAll : vpAll;
      const double Lb = viaMinus ? Lm : Lp;
      const double sgn = viaMinus ? -1.0 : 1.0;
      for (int q = 0; q < nc; ++q) {
        const double ph = oppOk ? phiOpp[q] : 0.0;
        const double Fb = pp ? (sgn * Lb) * ph : (((sgn * vb) * Lb) * dtn) * ph;
        red[q] = red[q] + Fb;
      }
    }
  }
  for (int q = 0; q < nc; ++q) Uout[q][c] = Uin[q][c] - red[q];

  // face fluxes for the dual remap (A1: mass, A2: volume), one face per
  // cell (its minus face) plus the last face of the line
  if (o.methode == 0 || o.methode == 1) {
    const int q0 = (o.methode == 0) ? m : 0;
    const long fm = XD ? j * (nx + 1) + a : a * nx + i;
    for (int k = 0; k < m; ++k) f.G[k][fm] = hasM ? Gm[q0 + k] : 0.0;
    if (!hasP) {
      const long fp = XD ? j * (nx + 1) + a + 1 : (a + 1) * nx + i;
      for (int k = 0; k < m; ++k) f.G[k][fp] = 0.0;
    }
  }
}

template __global__ void k_sweep<true>(KP);
template __global__ void k_sweep<false>(KP);

// Nodal momentum remap of one sweep (remap/DualPhaseRemap1.cc:18-136,
// remap/DualPhaseRemap2.cc:18-127 with remap/UtilesDualRemap.cc and
// UtilesRemap-Impl.h:8-97).  XD: along x (left/right nodes).
template <bool XD>
__global__ __launch_bounds__(256) void k_dual(KP kp) {
  const Opts& o = kp.o;
  const Fields& f = kp.f;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kp.nnode) return;
  const int m = o.nbmat;
  const long nx = o.nx, ny = o.ny;
  const long j = p / (nx + 1), i = p - j * (nx + 1);
  const long a = XD ? i : j, na = XD ? nx : ny;  // node along index in [0, na]
  const long stride = XD ? 1 : nx + 1;
  const double dtn = f.sc->dt[kp.cur];
  const double dt1 = f.sc->dt[kp.nxt];
  const int order2 = o.order - 1;

  // DualPhi and Lagrange coordinate of the nodes a-3..a+3
  double dpx[7], dpy[7], xl[7];
  bool in[7];
  for (int d = 0; d < 7; ++d) {
    const long b = a + d - 3;
    in[d] = (b >= 0 && b <= na);
    if (!in[d]) {
      dpx[d] = dpy[d] = xl[d] = 0.0;
      continue;
    }
    const long q = p + (d - 3) * stride;
    const double ux1 = f.ux[kp.nxt][q], uy1 = f.uy[kp.nxt][q];
    if (kp.sweep == 1) {
      dpx[d] = ux1, dpy[d] = uy1;
    } else {
      const double M1 = f.UD1[2][q];
      dpx[d] = f.UD1[0][q] / M1, dpy[d] = f.UD1[1][q] / M1;
    }
    const long jq = q / (nx + 1), iq = q - jq * (nx + 1);
    const D2 e = eulerNode(o, jq, iq);
    xl[d] = XD ? e.x + ux1 * dt1 : e.y + uy1 * dt1;
  }
  // limited nodal gradient at slot d (2..4) for velocity components 0, 1
  auto gradAt = [&](int d, double* g) {
    g[0] = g[1] = 0.0;
    if (o.order <= 1 || !in[d - 1] || !in[d + 1]) return;
    const double gr0 = (dpx[d + 1] - dpx[d]) / (xl[d + 1] - xl[d]);
    const double gl0 = (dpx[d] - dpx[d - 1]) / (xl[d] - xl[d - 1]);
    const double gr1 = (dpy[d + 1] - dpy[d]) / (xl[d + 1] - xl[d]);
    const double gl1 = (dpy[d] - dpy[d - 1]) / (xl[d] - xl[d - 1]);
    const double h0 = 0.5 * (xl[d + 1] - xl[d - 1]);
    double hm, hp;
    if (!in[d - 2]) {
      hm = 0., hp = 0.5 * (xl[d + 2] - xl[d]);
    } else if (!in[d + 2]) {
      hp = 0., hm = 0.5 * (xl[d] - xl[d - 2]);
    } else {
      hm = 0.5 * (xl[d] - xl[d - 2]);
      hp = 0.5 * (xl[d + 2] - xl[d]);
    }
    g[0] = limitedGradient(o.limiter, gr0, gl0, dpx[d], dpx[d + 1], dpx[d - 1], h0, hp, hm);
    g[1] = limitedGradient(o.limiter, gr1, gl1, dpy[d], dpy[d + 1], dpy[d - 1], h0, hp, hm);
  };
  double gM[2], g0[2], gP[2];
  gradAt(3, g0);
  gradAt(2, gM);
  gradAt(4, gP);

  // nodal mass fluxes through the dual faces on the minus / plus side
  double plus = 0., minus = 0.;
  {
    // the faces of the sweep orientation touching the node: two faces across
    // the sweep (t-1, t) at along index a, if that index is interior
    const long t = XD ? j : i, nt = XD ? ny : nx;
    const bool interior = (a > 0 && a < na);
    long tf[2];
    int nf = 0;
    if (interior) {
      if (t >= 1) tf[nf++] = t - 1;
      if (t <= nt - 1) tf[nf++] = t;
    }
    // G index of face (along b, across t); cells on both sides of it
    auto gidx = [&](long b, long tt) { return XD ? tt * (nx + 1) + b : b * nx + tt; };
    auto cidx = [&](long b, long tt) { return XD ? tt * nx + b : b * nx + tt; };
    for (int k = 0; k < m; ++k) {
      double pk = 0, mk = 0;
      if (o.methode == 0 || o.methode == 1) {
        for (int s = 0; s < nf; ++s) {
          const long tt = tf[s];
          const double Gf = f.G[k][gidx(a, tt)];
          double tp = Gf + f.G[k][gidx(a + 1, tt)];
          double tm = Gf + f.G[k][gidx(a - 1, tt)];
          if (o.methode == 1) tp = f.rL[cidx(a, tt)] * tp, tm = f.rL[cidx(a - 1, tt)] * tm;
          pk += tp;
          mk += tm;
        }
        const int nbfaces = 2 * nf;
        if (nbfaces != 0) pk /= nbfaces, mk /= nbfaces;
      } else if (o.methode == 2) {
        int nbcell = 0;
        for (int s = 0; s < nf; ++s) {
          const long tt = tf[s];
          pk += f.DF[k][cidx(a, tt)];
          mk += f.DF[k][cidx(a - 1, tt)];
          nbcell++;
        }
        if (nbcell != 0) pk /= nbcell, mk /= nbcell;
      }
      plus += pk;
      minus += mk;
    }
  }
  // upwind velocities (UtilesDualRemap.cc:234-397); a missing neighbour
  // reads as 0 and its mass flux is 0
  auto upwind = [&](bool minusSide, double flux, const double* gNb, double* r) {
    const int dn = minusSide ? 2 : 4;
    const double dNb[2] = {dpx[dn], dpy[dn]}, d0[2] = {dpx[3], dpy[3]};
    const double uN = XD ? dpx[dn] : dpy[dn], u0 = XD ? dpx[3] : dpy[3];
    const double u = minusSide ? 0.5 * (uN + u0) : 0.5 * (u0 + uN);
    const int sign = (u > 0) ? 1 : -1;
    const double h = minusSide ? (xl[3] - xl[dn]) : (xl[dn] - xl[3]);
    const bool fromNb = minusSide ? (flux > 0) : (flux < 0);
    for (int d = 0; d < 2; ++d) {
      const double base = fromNb ? dNb[d] : d0[d];
      const double g = fromNb ? gNb[d] : g0[d];
      r[d] = base + order2 * (0.5 * g * (sign * h - dtn * u));
    }
  };
  double uvm[2], uvp[2];
  upwind(true, minus, gM, uvm);
  upwind(false, plus, gP, uvp);
  double U0, U1, U2;
  if (kp.sweep == 1) {
    const double Mp = f.Mp[p];
    U0 = Mp * f.ux[kp.nxt][p], U1 = Mp * f.uy[kp.nxt][p], U2 = Mp;
  } else {
    U0 = f.UD1[0][p], U1 = f.UD1[1][p], U2 = f.UD1[2][p];
  }
  const double R2 = U2 + minus - plus;
  const double R0 = U0 + minus * uvm[0] - plus * uvp[0];
  const double R1 = U1 + minus * uvm[1] - plus * uvp[1];
  if (kp.sweep == 1) {
    f.UD1[0][p] = R0, f.UD1[1][p] = R1, f.UD1[2][p] = R2;
  } else {
    // remapVariables, nodes (PhaseRemapFinal.cc:207-234)
    f.ux[kp.nxt][p] = R0 / R2;
    f.uy[kp.nxt][p] = R1 / R2;
  }
}

template __global__ void k_dual<true>(KP);
template __global__ void k_dual<false>(KP);

// Vnr::remapVariables, cells (PhaseRemapFinal.cc:23-205), then computeEOS
// and computePressionMoyenne (PhaseLagrange.cc:392-485) on the remapped state.
__global__ __launch_bounds__(256) void k_remap_final(KP kp) {
  const Opts& o = kp.o;
  const Fields& f = kp.f;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= kp.ncell) return;
  const int m = o.nbmat;
  const int cur = kp.cur, nxt = kp.nxt;
  const double thr = o.threshold;
  double U[NCMAX];
  for (int q = 0; q < o.nc; ++q) U[q] = f.U2[q][c];
  const double vol = o.dx * o.dy;  // m_euler_volume (Init.cc:427)
  double volt = 0., masset = 0.;
  double vn[NMAT], fr[NMAT], mu[NMAT];
  for (int k = 0; k < m; ++k) {
    vn[k] = U[k];
    volt += vn[k];
    masset += U[m + k];
  }
  double vtn = 0.;
  for (int k = 0; k < m; ++k) {
    vn[k] *= vol / volt;
    vtn += vn[k];
  }
  double sf = 0.;
  for (int k = 0; k < m; ++k) {
    fr[k] = vn[k] / vtn;
    if (fr[k] < thr) fr[k] = 0.;
    sf += fr[k];
  }
  for (int k = 0; k < m; ++k) fr[k] = fr[k] / sf;
  int matcell = 0, imatpure = -1;
  for (int k = 0; k < m; ++k)
    if (fr[k] > 0.) matcell++, imatpure = k;
  const int mixte = (matcell > 1) ? 1 : 0;
  const int pure = (matcell > 1) ? -1 : imatpure;
  for (int k = 0; k < m; ++k) mu[k] = U[m + k] / masset;
  double fm = 0.;
  for (int k = 0; k < m; ++k) {
    if (fr[k] < thr) mu[k] = 0.;
    fm += mu[k];
  }
  for (int k = 0; k < m; ++k) mu[k] /= fm;
  double dens[NMAT], pe[NMAT];
  double dn = 0.;
  for (int k = 0; k < m; ++k) {
    dens[k] = 0.0;
    if (fr[k] > thr) dens[k] = U[m + k] / vn[k];
    dn += fr[k] * dens[k];
  }
  for (int k = 0; k < m; ++k) {
    pe[k] = 0.0;
    if ((fr[k] > thr) && (U[m + k] != 0.)) pe[k] = U[2 * m + k] / U[m + k];
  }
  const double rho1 = dn;
  double e1 = 0.;
  const double M = vol * rho1;
  double Mk[NMAT], rk[NMAT], ek[NMAT];
  for (int k = 0; k < m; ++k) Mk[k] = mu[k] * M;
  const double q1 = U[3 * m] / vol;
  for (int k = 0; k < m; ++k) {
    rk[k] = dens[k];
    ek[k] = pe[k];
    e1 += mu[k] * ek[k];
  }
  for (int k = 0; k < m; ++k) {
    if (pe[k] < 0. || dens[k] < 0.) rk[k] = 0., ek[k] = 0., mu[k] = 0., fr[k] = 0.;
  }
  if (e1 != e1 || rho1 != rho1) raise(f.sc, MHY_ERR_NAN);
  // computeEOS on the remapped state
  double pk[NMAT], ck[NMAT];
  for (int k = 0; k < m; ++k) {
    const double g = o.gammap[k];
    if (o.eos[k] == MHY_EOS_PERFECT_GAS) {
      pk[k] = (g - 1.0) * rk[k] * ek[k];
      ck[k] = sqrt(g * (g - 1.0) * ek[k]);
    } else if (o.eos[k] == MHY_EOS_VOID) {
      pk[k] = f.p_k[cur][k][c];
      ck[k] = f.c_k[cur][k][c];
      ek[k] = f.e_k[cur][k][c];
      e1 = 0.0;
    } else {
      pk[k] = f.p_k[nxt][k][c];
      ck[k] = f.c_k[nxt][k][c];
    }
  }
  // computePressionMoyenne
  double p1 = 0.0, c1 = 0.0;
  const bool needStale = !(rho1 > 0.);
  if (needStale) c1 = f.c[nxt][c];
  for (int k = 0; k < m; ++k) {
    p1 += fr[k] * pk[k];
    if (needStale) c1 = smax(c1, ck[k]);
  }
  if (rho1 > 0.) c1 = sqrt(o.gammap[0] * (o.gammap[0] - 1.0) * e1);
  for (int k = 0; k < m; ++k)
    if (o.eos[k] == MHY_EOS_VOID) e1 += mu[k] * ek[k];
  f.rho[nxt][c] = rho1;
  f.e[nxt][c] = e1;
  f.q[nxt][c] = q1;
  f.p[nxt][c] = p1;
  f.c[nxt][c] = c1;
  f.M[c] = M;
  f.mixte[c] = mixte;
  f.pure[c] = pure;
  for (int k = 0; k < m; ++k) {
    f.f_k[k][c] = fr[k];
    f.mu_k[k][c] = mu[k];
    f.M_k[k][c] = Mk[k];
    f.rho_k[nxt][k][c] = rk[k];
    f.e_k[nxt][k][c] = ek[k];
    f.p_k[nxt][k][c] = pk[k];
    f.c_k[nxt][k][c] = ck[k];
  }
}

// computeNodeMass for output (PhaseLagrange.cc:31-48)
__global__ void k_node_mass(KP kp) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kp.nnode) return;
  const long nx = kp.o.nx, ny = kp.o.ny;
  const long j = p / (nx + 1), i = p - j * (nx + 1);
  const bool hc[4] = {j < ny && i < nx, j < ny && i > 0, j > 0 && i < nx, j > 0 && i > 0};
  const long cc[4] = {j * nx + i, j * nx + i - 1, (j - 1) * nx + i, (j - 1) * nx + i - 1};
  double s = 0.0;
  int n = 0;
  for (int q = 0; q < 4; ++q)
    if (hc[q]) s = s + kp.f.M[cc[q]], ++n;
  kp.f.Mp[p] = s / (double)n;
}

// Conservation sums with Neumaier compensation: out[0] mass, out[1]
// internal energy (sum M*e), out[2+k] material masses.  Per-block partial
// pairs (sum, compensation) are combined on the host.
__global__ __launch_bounds__(256) void k_conservation(KP kp, double* out) {
  const Fields& f = kp.f;
  const int m = kp.o.nbmat;
  const int nq = 2 + m;
  double s[2 + NMAT], comp[2 + NMAT];
  for (int q = 0; q < nq; ++q) s[q] = comp[q] = 0.0;
  auto add = [&](int q, double v) {
    const double t = s[q] + v;
    if (fabs(s[q]) >= fabs(v)) comp[q] += (s[q] - t) + v;
    else comp[q] += (v - t) + s[q];
    s[q] = t;
  };
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < kp.ncell; c += (long)gridDim.x * blockDim.x) {
    const double M = f.M[c];
    add(0, M);
    add(1, M * f.e[kp.cur][c]);
    for (int k = 0; k < m; ++k) add(2 + k, f.M_k[k][c]);
  }
  __shared__ double sh[256];
  for (int q = 0; q < nq; ++q) {
    for (int pass = 0; pass < 2; ++pass) {
      sh[threadIdx.x] = pass == 0 ? s[q] : comp[q];
      __syncthreads();
      if (threadIdx.x == 0) {
        double t = 0.0, cc = 0.0;
        for (int k = 0; k < (int)blockDim.x; ++k) {
          const double v = sh[k];
          const double y = t + v;
          if (fabs(t) >= fabs(v)) cc += (t - y) + v;
          else cc += (v - y) + t;
          t = y;
        }
        out[(blockIdx.x * nq + q) * 4 + pass * 2] = t;
        out[(blockIdx.x * nq + q) * 4 + pass * 2 + 1] = cc;
      }
      __syncthreads();
    }
  }
}

}  // namespace mhy
