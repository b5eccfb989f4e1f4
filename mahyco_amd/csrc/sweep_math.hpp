// Per-cell arithmetic of one remap sweep.  M materials are a template
// parameter so the remap vectors (NC = 3M+1 components) live in registers;
// T is the arithmetic type of the slopes, limiters and PenteBorne profiles
// (double for the FP64 form, float for the mixed form), and every literal is
// a T so that the float form stays in FP32.  With T = double the operations
// are the reference's, in its order.
//
// Sources: remap/UtilesRemap.cc (limiters, PenteBorne helpers),
// remap/UtilesRemap-Impl.h (limited gradient, computeFluxPP[Pure], upwind
// face value), vnr_remap/PrepareRemap.cc + remap/PhaseRemap1.cc (Phi forms).
#pragma once

#include "common.hpp"

namespace mhy {
namespace sw {

// std::min / std::max argument conventions
template <class T>
__host__ __device__ __forceinline__ T tmin(T a, T b) {
  return (b < a) ? b : a;
}
template <class T>
__host__ __device__ __forceinline__ T tmax(T a, T b) {
  return (a < b) ? b : a;
}
template <class T>
__host__ __device__ __forceinline__ T tabs(T a) {
  return fabs(a);
}
// includes/Freefunctions.h:86-92
template <class T>
__host__ __device__ __forceinline__ T tdivide(T a, T b) {
  return (tabs(b) < T(1.0E-12)) ? T(0.0) : a / b;
}

// remap/UtilesRemap.cc:9-23
template <class T>
__host__ __device__ __forceinline__ T fluxLimiter(int lim, T r) {
  if (lim == MHY_LIM_MINMOD) return tmax(T(0.0), tmin(T(1.0), r));
  if (lim == MHY_LIM_SUPERBEE) return tmax(T(0.0), tmax(tmin(T(2.0) * r, T(1.0)), tmin(r, T(2.0))));
  if (lim == MHY_LIM_VANLEER) return (r <= T(0.0)) ? T(0.0) : T(2.0) * r / (T(1.0) + r);
  return T(0.0);
}

// remap/UtilesRemap.cc:25-72
template <class T>
__host__ __device__ __forceinline__ T fluxLimiterPP(int lim, T gradplus, T gradmoins, T yplus, T ymoins, T h0,
                                                    T hplus, T hmoins) {
  T grady = T(0.0);
  if (gradplus * gradmoins < T(0.0)) return T(0.);
  if (lim == MHY_LIM_MINMODG) {
    grady = ((yplus - ymoins) > T(0.)) ? tmin(tabs(gradplus), tabs(gradmoins)) : -tmin(tabs(gradplus), tabs(gradmoins));
  } else if (lim == MHY_LIM_SUPERBEEG) {
    grady = ((yplus - ymoins) > T(0.)) ? tmax(tabs(gradplus), tabs(gradmoins)) : -tmax(tabs(gradplus), tabs(gradmoins));
  } else if (lim == MHY_LIM_VANLEERG) {
    T lp = (h0 / T(2.) + hplus) / (h0 + hplus + hmoins);
    T lm = (h0 / T(2.) + hmoins) / (h0 + hplus + hmoins);
    if ((lp * gradplus + lm * gradmoins) != T(0.)) grady = gradplus * gradmoins / (lp * gradplus + lm * gradmoins);
    else grady = T(0.);
  } else if (lim == MHY_LIM_ULTRABEEG) {
    grady = (yplus - ymoins) / h0;
  } else if (lim == MHY_LIM_ARITHMETICG) {
    T lp = (h0 / T(2.) + hplus) / (h0 + hplus + hmoins);
    T lm = (h0 / T(2.) + hmoins) / (h0 + hplus + hmoins);
    grady = lm * gradplus + lp * gradmoins;
  }
  T gMp = gradplus * (h0 + hplus) / h0;
  T gMm = gradmoins * (h0 + hmoins) / h0;
  T gM = tmin(tabs(gMp), tabs(gMm));
  if ((yplus - ymoins) > T(0.)) grady = tmin(tabs(gM), tabs(grady));
  else grady = -tmin(tabs(gM), tabs(grady));
  return grady;
}

// remap/UtilesRemap-Impl.h:99-126, one component:
//   (limiter(tdivide(gp, gm)) * gm + limiter(tdivide(gm, gp)) * gp) / 2.
// With minmod one quotient suffices.  The quotient whose numerator has the
// larger magnitude is, correctly rounded, at least 1 in magnitude, where
// minmod gives exactly 1 for a positive and exactly +0 for a negative
// argument (and tdivide gives 0 when the other operand is below 1e-12), so
// only the other quotient is formed.  Both limiter values, and so the
// result, are bitwise those of the two-quotient form, NaN and infinite
// operands included (the result is then NaN in both forms).
template <class T>
__host__ __device__ __forceinline__ T limitedGradient(int lim, T gp, T gm, T phip, T phim, T h0, T hp, T hm) {
  if (lim == MHY_LIM_MINMOD) {
    const bool pSmall = tabs(gp) < tabs(gm);
    const T num = pSmall ? gp : gm, den = pSmall ? gm : gp;
    const T q = fluxLimiter<T>(lim, tdivide(num, den));  // |num| <= |den| (or a NaN)
    const T s = (tabs(num) < T(1.0E-12) || ((num > T(0.0)) != (den > T(0.0)))) ? T(0.0) : T(1.0);
    const T lp = pSmall ? q : s, lm = pSmall ? s : q;  // limiter(gp/gm), limiter(gm/gp)
    return (lp * gm + lm * gp) / T(2.0);
  }
  if (lim < MHY_LIM_MINMODG)
    return (fluxLimiter<T>(lim, tdivide(gp, gm)) * gm + fluxLimiter<T>(lim, tdivide(gm, gp)) * gp) / T(2.0);
  return fluxLimiterPP<T>(lim, gp, gm, phip, phim, h0, hp, hm);
}

// remap/UtilesRemap.cc:74-125 (both thresholds)
template <class T>
__device__ __forceinline__ void computeY0(int lim, T yplus, T ymoins, T h0, T hplus, T hmoins, T& y0plus, T& y0moins) {
  y0plus = T(0.), y0moins = T(0.);
  if (lim == MHY_LIM_MINMODG || lim == MHY_LIM_MINMOD || lim == 3000) {
    y0plus = yplus, y0moins = ymoins;
  } else if (lim == MHY_LIM_SUPERBEEG || lim == MHY_LIM_SUPERBEE) {
    y0plus = ((h0 + hmoins) * yplus + h0 * ymoins) / (T(2) * h0 + hmoins);
    y0moins = ((h0 + hplus) * ymoins + h0 * yplus) / (T(2) * h0 + hplus);
  } else if (lim == MHY_LIM_VANLEERG || lim == MHY_LIM_VANLEER) {
    T a = tmin(yplus, ymoins), b = tmax(yplus, ymoins);
    T xplus = (h0 * h0 + T(3) * h0 * hmoins + T(2) * hmoins * hmoins) * yplus;
    T xmoins = (h0 * h0 + T(3) * h0 * hplus + T(2) * hplus * hplus) * ymoins;
    xplus += (h0 * h0 - h0 * hplus - T(2) * hplus * hplus + T(2) * h0 * hmoins) * ymoins;
    xmoins += (h0 * h0 - h0 * hmoins - T(2) * hmoins * hmoins + T(2) * h0 * hplus) * yplus;
    xplus /= (T(2) * h0 * h0 + T(5) * h0 * hmoins + T(2) * hmoins * hmoins - h0 * hplus - T(2) * hplus * hplus);
    xmoins /= (T(2) * h0 * h0 + T(5) * h0 * hplus + T(2) * hplus * hplus - h0 * hmoins - T(2) * hmoins * hmoins);
    y0plus = tmin(tmax(xplus, a), b);
    y0moins = tmin(tmax(xmoins, a), b);
  } else if (lim == MHY_LIM_ULTRABEEG) {
    y0plus = (yplus + ymoins) / T(2.);
    y0moins = (yplus + ymoins) / T(2.);
  } else if (lim == MHY_LIM_ARITHMETICG) {
    y0plus = ((h0 + hmoins + hplus) * yplus + h0 * ymoins) / (T(2) * h0 + hmoins + hplus);
    y0moins = ((h0 + hmoins + hplus) * ymoins + h0 * yplus) / (T(2) * h0 + hmoins + hplus);
  }
}

// remap/UtilesRemap.cc:177-188
template <class T>
__device__ __forceinline__ T INT2Y(T X, T x0, T y0, T x1, T y1) {
  T flux = T(0.);
  if (tabs(x1 - x0) > T(1.e-14)) {
    T eta = tmin(tmax(T(0.), (X - x0) / (x1 - x0)), T(1.));
    flux = (y0 + T(0.5) * eta * (y1 - y0)) * (x1 - x0) * eta;
  }
  return flux;
}

// Sub-cell profile of one component (UtilesRemap.cc:127-166)
template <class T>
struct Profile {
  T xg, xd, yg, yd;
};
template <class T>
__device__ __forceinline__ Profile<T> profile(int lim, T thr, T y0, T yplus, T ymoins, T h0, T hplus, T hmoins,
                                              T grady) {
  T y0p, y0m;
  computeY0<T>(lim, yplus, ymoins, h0, hplus, hmoins, y0p, y0m);
  Profile<T> pr;
  T xplus = T(1.);
  if (tabs(y0p - yplus) > thr) xplus = (y0 - yplus) / (y0p - yplus) - T(1.) / T(2.);
  T xmoins = T(1.);
  if (tabs(y0m - ymoins) > thr) xmoins = (y0 - ymoins) / (y0m - ymoins) - T(1.) / T(2.);
  pr.xd = +h0 * tmin(tmax(xplus, T(-1.) / T(2.)), T(1.) / T(2.));
  pr.xg = -h0 * tmin(tmax(xmoins, T(-1.) / T(2.)), T(1.) / T(2.));
  const T xtd = y0 + h0 / T(2) * grady;
  const T xtg = y0 - h0 / T(2) * grady;
  const T ad = tmin(yplus, T(2.) * y0m - ymoins), bd = tmax(yplus, T(2.) * y0m - ymoins);
  const T ag = tmin(ymoins, T(2.) * y0p - yplus), bg = tmax(ymoins, T(2.) * y0p - yplus);
  pr.yd = tmin(tmax(xtd, ad), bd);
  pr.yg = tmin(tmax(xtg, ag), bg);
  return pr;
}

// swept integral of the three-segment profile between lo and hi
// (UtilesRemap-Impl.h:185-200, :230-247)
template <class T>
__device__ __forceinline__ T sweptIntegral(const Profile<T>& pr, T h0, T ymoins, T yplus, T lo, T hi) {
  const T f1b = INT2Y<T>(hi, -h0 / T(2.), ymoins, pr.xg, pr.yg), f1a = INT2Y<T>(lo, -h0 / T(2.), ymoins, pr.xg, pr.yg);
  const T f2b = INT2Y<T>(hi, pr.xg, pr.yg, pr.xd, pr.yd), f2a = INT2Y<T>(lo, pr.xg, pr.yg, pr.xd, pr.yd);
  const T f3b = INT2Y<T>(hi, pr.xd, pr.yd, h0 / T(2.), yplus), f3a = INT2Y<T>(lo, pr.xd, pr.yd, h0 / T(2.), yplus);
  return tmax(((f1b - f1a) + (f2b - f2a) + (f3b - f3a)), T(0.));
}

// Phi of a cell at the start of the sweep: U/V, or the PenteBorne form
// (fractions, densities, specific energies, then the Lagrange-phase pseudo
// of VNR or the velocity and specific kinetic energy of Eucclhyd: the
// momentum components over the cell mass, PhaseRemap1.cc:530-547,
// eucclhyd_remap/PhaseLagrange.cc:932-958), formed in FP64 from the FP64
// remap vector and then held as T; an FP32 U/V (mixed form, limited slopes)
// is formed in FP32.  L: the vector layout (kMat, kNC).
template <int L, bool PP, class T>
__device__ __forceinline__ void cellPhi(const double* u, double V, double qL, T* phi) {
  constexpr int NC = kNC<L>;
  constexpr int M = kMat<L>;
  if (PP) {
    double sv = 0.0, sm = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) sv += u[k];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      phi[k] = T(u[k] / sv);
      phi[M + k] = T((u[k] != 0.) ? u[M + k] / u[k] : 0.0);
      sm += u[M + k];
    }
#pragma unroll
    for (int k = 0; k < M; ++k) phi[2 * M + k] = T((u[M + k] != 0.) ? u[2 * M + k] / u[M + k] : 0.0);
    if constexpr (kEuc<L>) {
#pragma unroll
      for (int q = 3 * M; q < NC; ++q) phi[q] = T(u[q] / sm);
    } else {
      phi[3 * M] = T(qL);
    }
  } else if (sizeof(T) == 4) {
    // FP32 slopes: Phi = U * (1/V) in FP32 (one reciprocal per cell)
    const T rV = T(1.0) / T(V);
#pragma unroll
    for (int q = 0; q < NC; ++q) phi[q] = T(u[q]) * rV;
  } else {
#pragma unroll
    for (int q = 0; q < NC; ++q) phi[q] = T(u[q] / V);
  }
}

// mixed / pure indicators from a Phi vector (PhaseRemap1.cc:552-566)
template <int L, class T>
__device__ __forceinline__ int packedFlags(const T* phi) {
  int matcell = 0, imatpure = -1;
#pragma unroll
  for (int k = 0; k < kMat<L>; ++k)
    if (phi[k] > T(0.)) matcell++, imatpure = k;
  const int mixte = (matcell > 1) ? 1 : 0;
  const int pure = (matcell > 1) ? -1 : imatpure;
  return (mixte << 8) | (pure + 1);
}
__device__ __forceinline__ int packFlags(int mixte, int pure) { return (mixte << 8) | (pure + 1); }

// Inputs of one cell's slope / PenteBorne computation: itself and the
// neighbours along the sweep, clamped to the cell at the domain edge.
template <class T>
struct Nbr {
  const T *phim, *phi, *phip;  // Phi of minus neighbour, cell, plus neighbour
  T xcm, xc, xcp;              // along-sweep Lagrange centroid coordinates
  bool inM, inP;               // the minus / plus neighbour exists (inner face)
  T h0, hm, hp;                // widths (PenteBorne and generalised limiters)
  bool vp;                     // pure neighbourhood (PenteBorne mixte option)
  // divided differences at the minus / plus faces when the caller has them
  // (a neighbour's minus face is this cell's plus face); null: formed here
  const T *dM = nullptr, *dP = nullptr;
};

// Divided differences at the minus / plus faces (gradPhiFace, PhaseRemap1.cc:
// 39-40, 87-88): (Phi(front) - Phi(back)) / deltaxLagrange; x: back = left,
// y: back = above, so the y form is written below-minus-above.  Faces on the
// domain edge keep their zero initial value.
template <bool XD, class T>
__device__ __forceinline__ T faceDiffM(const Nbr<T>& n, int q) {
  if (!n.inM) return T(0.0);
  return XD ? (n.phi[q] - n.phim[q]) / (n.xc - n.xcm) : (n.phim[q] - n.phi[q]) / (n.xcm - n.xc);
}
template <bool XD, class T>
__device__ __forceinline__ T faceDiffP(const Nbr<T>& n, int q) {
  if (!n.inP) return T(0.0);
  return XD ? (n.phip[q] - n.phi[q]) / (n.xcp - n.xc) : (n.phi[q] - n.phip[q]) / (n.xc - n.xcp);
}

// Number of leading components whose PenteBorne swept integral is used:
// with the complete option every component (the Eucclhyd momentum and
// kinetic energy are rescaled from their own integrals), else the volumes,
// and the masses too in a pure neighbourhood.
template <int L>
__device__ __forceinline__ int ppIntegrated(const Opts& o, bool vp) {
  constexpr int M = kMat<L>;
  return (o.pp_complet == 1) ? (kEuc<L> ? kNC<L> : 3 * M) : (vp ? 2 * M : M);
}

// Finishing of a PenteBorne flux vector (UtilesRemap-Impl.h:259-313,
// :436-487): masses and energies from the volume fluxes, the VNR
// pseudo-viscosity with the volume flux, the Eucclhyd momentum and kinetic
// energy with the mass flux.
template <int L, class T>
__device__ __forceinline__ void finishPP(const Opts& o, bool pureVariant, const T* phi, T* F) {
  constexpr int NC = kNC<L>;
  constexpr int M = kMat<L>;
  if (o.pp_complet == 1) {
    T sfv = T(0.), sfm = T(0.);
#pragma unroll
    for (int k = 0; k < M; ++k) sfv += F[k];
    if (tabs(sfv) > T(o.threshold)) {
#pragma unroll
      for (int k = 0; k < M; ++k) {
        if (!pureVariant) F[M + k] = (F[M + k] / sfv) * F[k];
        F[2 * M + k] = (F[2 * M + k] / sfv) * F[M + k];
        sfm += F[M + k];
      }
      if constexpr (kEuc<L>) {
#pragma unroll
        for (int q = 3 * M; q < NC; ++q) F[q] = (F[q] / sfv) * sfm;
      } else {
        F[3 * M] = phi[3 * M] * sfv;
      }
    } else {
#pragma unroll
      for (int q = 0; q < NC; ++q) F[q] = T(0.0);
    }
  } else {
    T sfv = T(0.), sfm = T(0.);
#pragma unroll
    for (int k = 0; k < M; ++k) {
      if (!pureVariant) F[M + k] = phi[M + k] * F[k];
      F[2 * M + k] = phi[2 * M + k] * F[M + k];
      sfm += F[M + k];
      sfv += F[k];
    }
    if constexpr (kEuc<L>) {
#pragma unroll
      for (int q = 3 * M; q < NC; ++q) F[q] = phi[q] * sfm;
    } else {
      F[3 * M] = phi[3 * M] * sfv;
    }
  }
}

// Limited gradient of component q (XD: gradplus is the plus face; along y the
// reference passes the bottom (minus) face as gradplus, PhaseRemap1.cc:262-265).
template <bool XD, class T>
__device__ __forceinline__ T gradQ(const Opts& o, const Nbr<T>& n, int q) {
  if (o.order <= 1) return T(0.0);  // gradients are only formed for order > 1
  const T gM = n.dM ? n.dM[q] : faceDiffM<XD, T>(n, q);
  const T gP = n.dP ? n.dP[q] : faceDiffP<XD, T>(n, q);
  const int lim = (o.pp == 1 && n.vp) ? o.limiter_pure : o.limiter;
  const T gp = XD ? gP : gM, gm = XD ? gM : gP;
  return limitedGradient<T>(lim, gp, gm, n.phip[q], n.phim[q], n.h0, n.hp, n.hm);
}

// PenteBorne fluxes of one cell through its minus face (Ar, type 0) and its
// plus face (Av, type 1), and the dual flux (PB) of the mass components.
// vminus / vplus: normal velocities of the minus / plus faces.  WANT_AR /
// WANT_AV select the face fluxes the caller uses (the other stays zero).
template <int L, bool XD, class T, bool WANT_AR = true, bool WANT_AV = true>
__device__ __forceinline__ void ppFluxes(const Opts& o, const Nbr<T>& n, T vminus, T vplus, T dtn, bool dual, T* Ar,
                                         T* Av, T* Dual) {
  constexpr int NC = kNC<L>;
  constexpr int M = kMat<L>;
  const int lz = n.vp ? o.limiter_pure : o.limiter;
  const T h0 = n.h0, hp = n.hp, hm = n.hm;
  const T thr = T(o.threshold);
  const T ar = -vminus;  // outer normal . e = -1 on the minus face
  const T av = vplus;
  const T dualv = T(0.5) * (vminus + vplus);
  const T ppvR = T(0.5) * (ar + tabs(ar)) * dtn;
  const T ppvV = T(0.5) * (av + tabs(av)) * dtn;
  const T ppdv = T(0.5) * (dualv + tabs(dualv)) * dtn;
#pragma unroll
  for (int q = 0; q < NC; ++q) Ar[q] = Av[q] = T(0.0);
#pragma unroll
  for (int k = 0; k < M; ++k) Dual[k] = T(0.0);
  const int nInt = ppIntegrated<L>(o, n.vp);
#pragma unroll
  for (int q = 0; q < (kEuc<L> ? NC : 3 * M); ++q) {
    const bool needF = (WANT_AR || WANT_AV) && q < nInt;
    const bool needD = dual && q >= M && q < 2 * M;
    if (!needF && !needD) continue;
    const T phi = n.phi[q], phip = n.phip[q], phim = n.phim[q];
    const T g = gradQ<XD, T>(o, n, q);
    const Profile<T> pr = profile<T>(lz, thr, phi, phip, phim, h0, hp, hm, g);
    const bool flat = ((phip - phi) * (phim - phi)) >= T(0.);
    if (needF) {
      if (WANT_AR) Ar[q] = flat ? phi * ppvR : sweptIntegral<T>(pr, h0, phim, phip, -h0 / T(2.), -h0 / T(2.) + ppvR);
      if (WANT_AV) Av[q] = flat ? phi * ppvV : sweptIntegral<T>(pr, h0, phim, phip, h0 / T(2.) - ppvV, h0 / T(2.));
    }
    if (needD) Dual[q - M] = flat ? phi * ppdv : sweptIntegral<T>(pr, h0, phim, phip, T(0.), ppdv);
  }
  if (WANT_AR) {
    finishPP<L, T>(o, n.vp, n.phi, Ar);
    if (ppvR == T(0.))
#pragma unroll
      for (int q = 0; q < NC; ++q) Ar[q] = T(0.0);
  }
  if (WANT_AV) {
    finishPP<L, T>(o, n.vp, n.phi, Av);
    if (ppvV == T(0.))
#pragma unroll
      for (int q = 0; q < NC; ++q) Av[q] = T(0.0);
  }
  if (ppdv == T(0.))
#pragma unroll
    for (int k = 0; k < M; ++k) Dual[k] = T(0.0);
}

}  // namespace sw
}  // namespace mhy
