// Per-cell arithmetic of one remap sweep, shared by the x (LDS tile) and y
// (column walker) kernels.  M materials are a template parameter so the
// remap vectors (NC = 3M+1 components) live in registers.
//
// Sources: remap/UtilesRemap.cc (limiters, PenteBorne helpers),
// remap/UtilesRemap-Impl.h (limited gradient, computeFluxPP[Pure], upwind
// face value), vnr_remap/PrepareRemap.cc + remap/PhaseRemap1.cc (Phi forms).
#pragma once

#include "common.hpp"

namespace mhy {
namespace sw {

// remap/UtilesRemap.cc:9-23
__device__ __forceinline__ double fluxLimiter(int lim, double r) {
  if (lim == MHY_LIM_MINMOD) return smax(0.0, smin(1.0, r));
  if (lim == MHY_LIM_SUPERBEE) return smax(0.0, smax(smin(2.0 * r, 1.0), smin(r, 2.0)));
  if (lim == MHY_LIM_VANLEER) return (r <= 0.0) ? 0.0 : 2.0 * r / (1.0 + r);
  return 0.0;
}

// remap/UtilesRemap.cc:25-72
__device__ __forceinline__ double fluxLimiterPP(int lim, double gradplus, double gradmoins, double yplus, double ymoins,
                                                double h0, double hplus, double hmoins) {
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
__device__ __forceinline__ double limitedGradient(int lim, double gp, double gm, double phip, double phim, double h0,
                                                  double hp, double hm) {
  if (lim < MHY_LIM_MINMODG)
    return (fluxLimiter(lim, divideNoExcept(gp, gm)) * gm + fluxLimiter(lim, divideNoExcept(gm, gp)) * gp) / 2.0;
  return fluxLimiterPP(lim, gp, gm, phip, phim, h0, hp, hm);
}

// remap/UtilesRemap.cc:74-125 (both thresholds)
__device__ __forceinline__ void computeY0(int lim, double yplus, double ymoins, double h0, double hplus, double hmoins,
                                          double& y0plus, double& y0moins) {
  y0plus = 0., y0moins = 0.;
  if (lim == MHY_LIM_MINMODG || lim == MHY_LIM_MINMOD || lim == 3000) {
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

// Sub-cell profile of one component (UtilesRemap.cc:127-166)
struct Profile {
  double xg, xd, yg, yd;
};
__device__ __forceinline__ Profile profile(int lim, double thr, double y0, double yplus, double ymoins, double h0,
                                           double hplus, double hmoins, double grady) {
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

// swept integral of the three-segment profile between lo and hi
// (UtilesRemap-Impl.h:185-200, :230-247)
__device__ __forceinline__ double sweptIntegral(const Profile& pr, double h0, double ymoins, double yplus, double lo,
                                                double hi) {
  const double f1b = INT2Y(hi, -h0 / 2., ymoins, pr.xg, pr.yg), f1a = INT2Y(lo, -h0 / 2., ymoins, pr.xg, pr.yg);
  const double f2b = INT2Y(hi, pr.xg, pr.yg, pr.xd, pr.yd), f2a = INT2Y(lo, pr.xg, pr.yg, pr.xd, pr.yd);
  const double f3b = INT2Y(hi, pr.xd, pr.yd, h0 / 2., yplus), f3a = INT2Y(lo, pr.xd, pr.yd, h0 / 2., yplus);
  return smax(((f1b - f1a) + (f2b - f2a) + (f3b - f3a)), 0.);
}

// Phi of a cell at the start of the sweep: U/V, or the PenteBorne form
// (fractions, densities, specific energies, Lagrange-phase pseudo).
template <int M, bool PP>
__device__ __forceinline__ void cellPhi(const double* u, double V, double qL, double* phi) {
  constexpr int NC = 3 * M + 1;
  if (PP) {
    double sv = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) sv += u[k];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      phi[k] = u[k] / sv;
      phi[M + k] = (u[k] != 0.) ? u[M + k] / u[k] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < M; ++k) phi[2 * M + k] = (u[M + k] != 0.) ? u[2 * M + k] / u[M + k] : 0.0;
    phi[3 * M] = qL;
  } else {
#pragma unroll
    for (int q = 0; q < NC; ++q) phi[q] = u[q] / V;
  }
}

// mixed / pure indicators from a Phi vector (PhaseRemap1.cc:552-566)
template <int M>
__device__ __forceinline__ int packedFlags(const double* phi) {
  int matcell = 0, imatpure = -1;
#pragma unroll
  for (int k = 0; k < M; ++k)
    if (phi[k] > 0.) matcell++, imatpure = k;
  const int mixte = (matcell > 1) ? 1 : 0;
  const int pure = (matcell > 1) ? -1 : imatpure;
  return (mixte << 8) | (pure + 1);
}
__device__ __forceinline__ int packFlags(int mixte, int pure) { return (mixte << 8) | (pure + 1); }

// Inputs of one cell's slope / PenteBorne computation: itself and the
// neighbours along the sweep, clamped to the cell at the domain edge.
struct Nbr {
  const double *phim, *phi, *phip;  // Phi of minus neighbour, cell, plus neighbour
  double xcm, xc, xcp;              // along-sweep Lagrange centroid coordinates
  bool inM, inP;                    // the minus / plus neighbour exists (inner face)
  double h0, hm, hp;                // widths (PenteBorne and generalised limiters)
  bool vp;                          // pure neighbourhood (PenteBorne mixte option)
};

// Divided differences at the minus / plus faces (gradPhiFace, PhaseRemap1.cc:
// 39-40, 87-88): (Phi(front) - Phi(back)) / deltaxLagrange; x: back = left,
// y: back = above, so the y form is written below-minus-above.  Faces on the
// domain edge keep their zero initial value.
template <bool XD>
__device__ __forceinline__ double faceDiffM(const Nbr& n, int q) {
  if (!n.inM) return 0.0;
  return XD ? (n.phi[q] - n.phim[q]) / (n.xc - n.xcm) : (n.phim[q] - n.phi[q]) / (n.xcm - n.xc);
}
template <bool XD>
__device__ __forceinline__ double faceDiffP(const Nbr& n, int q) {
  if (!n.inP) return 0.0;
  return XD ? (n.phip[q] - n.phi[q]) / (n.xcp - n.xc) : (n.phi[q] - n.phip[q]) / (n.xc - n.xcp);
}

// Number of leading components whose PenteBorne swept integral is used.
template <int M>
__device__ __forceinline__ int ppIntegrated(const Opts& o, bool vp) {
  return (o.pp_complet == 1) ? 3 * M : (vp ? 2 * M : M);
}

// Finishing of a PenteBorne flux vector (UtilesRemap-Impl.h:259-313, :436-487)
template <int M>
__device__ __forceinline__ void finishPP(const Opts& o, bool pureVariant, const double* phi, double* F) {
  constexpr int NC = 3 * M + 1;
  if (o.pp_complet == 1) {
    double sfv = 0.;
#pragma unroll
    for (int k = 0; k < M; ++k) sfv += F[k];
    if (fabs(sfv) > o.threshold) {
#pragma unroll
      for (int k = 0; k < M; ++k) {
        if (!pureVariant) F[M + k] = (F[M + k] / sfv) * F[k];
        F[2 * M + k] = (F[2 * M + k] / sfv) * F[M + k];
      }
      F[3 * M] = phi[3 * M] * sfv;
    } else {
#pragma unroll
      for (int q = 0; q < NC; ++q) F[q] = 0.0;
    }
  } else {
    double sfv = 0.;
#pragma unroll
    for (int k = 0; k < M; ++k) {
      if (!pureVariant) F[M + k] = phi[M + k] * F[k];
      F[2 * M + k] = phi[2 * M + k] * F[M + k];
      sfv += F[k];
    }
    F[3 * M] = phi[3 * M] * sfv;
  }
}

// Limited gradient of component q (XD: gradplus is the plus face; along y the
// reference passes the bottom (minus) face as gradplus, PhaseRemap1.cc:262-265).
template <bool XD>
__device__ __forceinline__ double gradQ(const Opts& o, const Nbr& n, int q) {
  if (o.order <= 1) return 0.0;  // gradients are only formed for order > 1
  const double gM = faceDiffM<XD>(n, q), gP = faceDiffP<XD>(n, q);
  const int lim = (o.pp == 1 && n.vp) ? o.limiter_pure : o.limiter;
  const double gp = XD ? gP : gM, gm = XD ? gM : gP;
  return limitedGradient(lim, gp, gm, n.phip[q], n.phim[q], n.h0, n.hp, n.hm);
}

// PenteBorne fluxes of one cell through its minus face (Ar, type 0) and its
// plus face (Av, type 1), and the dual flux (PB) of the mass components.
// vminus / vplus: normal velocities of the minus / plus faces.  WANT_AR /
// WANT_AV select the face fluxes the caller uses (the other stays zero).
template <int M, bool XD, bool WANT_AR = true, bool WANT_AV = true>
__device__ __forceinline__ void ppFluxes(const Opts& o, const Nbr& n, double vminus, double vplus, double dtn,
                                         bool dual, double* Ar, double* Av, double* Dual) {
  constexpr int NC = 3 * M + 1;
  const int lz = n.vp ? o.limiter_pure : o.limiter;
  const double h0 = n.h0, hp = n.hp, hm = n.hm;
  const double ar = -vminus;  // outer normal . e = -1 on the minus face
  const double av = vplus;
  const double dualv = 0.5 * (vminus + vplus);
  const double ppvR = 0.5 * (ar + fabs(ar)) * dtn;
  const double ppvV = 0.5 * (av + fabs(av)) * dtn;
  const double ppdv = 0.5 * (dualv + fabs(dualv)) * dtn;
#pragma unroll
  for (int q = 0; q < NC; ++q) Ar[q] = Av[q] = 0.0;
#pragma unroll
  for (int k = 0; k < M; ++k) Dual[k] = 0.0;
  const int nInt = ppIntegrated<M>(o, n.vp);
#pragma unroll
  for (int q = 0; q < 3 * M; ++q) {
    const bool needF = (WANT_AR || WANT_AV) && q < nInt;
    const bool needD = dual && q >= M && q < 2 * M;
    if (!needF && !needD) continue;
    const double phi = n.phi[q], phip = n.phip[q], phim = n.phim[q];
    const double g = gradQ<XD>(o, n, q);
    const Profile pr = profile(lz, o.threshold, phi, phip, phim, h0, hp, hm, g);
    const bool flat = ((phip - phi) * (phim - phi)) >= 0.;
    if (needF) {
      if (WANT_AR) Ar[q] = flat ? phi * ppvR : sweptIntegral(pr, h0, phim, phip, -h0 / 2., -h0 / 2. + ppvR);
      if (WANT_AV) Av[q] = flat ? phi * ppvV : sweptIntegral(pr, h0, phim, phip, h0 / 2. - ppvV, h0 / 2.);
    }
    if (needD) Dual[q - M] = flat ? phi * ppdv : sweptIntegral(pr, h0, phim, phip, 0., ppdv);
  }
  if (WANT_AR) {
    finishPP<M>(o, n.vp, n.phi, Ar);
    if (ppvR == 0.)
#pragma unroll
      for (int q = 0; q < NC; ++q) Ar[q] = 0.0;
  }
  if (WANT_AV) {
    finishPP<M>(o, n.vp, n.phi, Av);
    if (ppvV == 0.)
#pragma unroll
      for (int q = 0; q < NC; ++q) Av[q] = 0.0;
  }
  if (ppdv == 0.)
#pragma unroll
    for (int k = 0; k < M; ++k) Dual[k] = 0.0;
}

}  // namespace sw
}  // namespace mhy
