// Final remap of one cell (Vnr::remapVariables, cells: PhaseRemapFinal.cc:23-205)
// followed by computeEOS and computePressionMoyenne (PhaseLagrange.cc:392-485)
// on the remapped state.  Shared by the stand-alone final-remap kernel and by
// the second remap sweep, which finishes each cell from its registers.
#pragma once

#include "common.hpp"

namespace mhy {

// first error wins (MHY_ERR_*), read by the host after the step
__device__ __forceinline__ void raiseErr(Scalars* sc, int code) { atomicCAS(&sc->err, 0, code); }

// What the final remap stores for one cell, already in the storage types.
template <class R, int M>
struct FinalOut {
  R rho, e, q, p, c;
  int mixte, pure;
  double Mc;
  double fr[M], Mk[M];
  R mu[M], rk[M], ek[M], pk[M], ck[M];
};

// U: the 3m+1 live components of the remapped vector U2 of cell c.
// The volume fractions near 0 and 1, their threshold tests, the densities,
// specific energies and masses are FP64 in both forms (a nearly emptied
// material's mass and energy content can lie below the FP32 range).  The
// mixed form (R = float) takes the EOS sound speeds' square roots in FP32,
// the precision they are stored in.
template <class R, int M>
__device__ __forceinline__ void finalCompute(const KPT<R>& kp, long c, const double* U, FinalOut<R, M>& out) {
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const int cur = kp.cur, nxt = kp.nxt;
  const double thr = o.threshold;
  const double vol = o.dx * o.dy;  // m_euler_volume (Init.cc:427)
  double volt = 0., masset = 0.;
  double vn[M], fr[M], mu[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    vn[k] = U[k];
    volt += vn[k];
    masset += U[M + k];
  }
  double vtn = 0.;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    vn[k] *= vol / volt;
    vtn += vn[k];
  }
  double sf = 0.;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    fr[k] = vn[k] / vtn;
    if (fr[k] < thr) fr[k] = 0.;
    sf += fr[k];
  }
#pragma unroll
  for (int k = 0; k < M; ++k) fr[k] = fr[k] / sf;
  int matcell = 0, imatpure = -1;
#pragma unroll
  for (int k = 0; k < M; ++k)
    if (fr[k] > 0.) matcell++, imatpure = k;
  const int mixte = (matcell > 1) ? 1 : 0;
  const int pure = (matcell > 1) ? -1 : imatpure;
#pragma unroll
  for (int k = 0; k < M; ++k) mu[k] = U[M + k] / masset;
  double fm = 0.;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    if (fr[k] < thr) mu[k] = 0.;
    fm += mu[k];
  }
#pragma unroll
  for (int k = 0; k < M; ++k) mu[k] /= fm;
  double dens[M], pe[M];
  double dn = 0.;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    dens[k] = 0.0;
    if (fr[k] > thr) dens[k] = U[M + k] / vn[k];
    dn += fr[k] * dens[k];
  }
#pragma unroll
  for (int k = 0; k < M; ++k) {
    pe[k] = 0.0;
    if ((fr[k] > thr) && (U[M + k] != 0.))
      pe[k] = U[2 * M + k] / U[M + k];
  }
  const double rho1 = dn;
  double e1 = 0.;
  const double Mc = vol * rho1;
  double Mk[M], rk[M], ek[M];
#pragma unroll
  for (int k = 0; k < M; ++k) Mk[k] = mu[k] * Mc;
  const double q1 = U[3 * M] / vol;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    rk[k] = dens[k];
    ek[k] = pe[k];
    e1 += mu[k] * ek[k];
  }
#pragma unroll
  for (int k = 0; k < M; ++k) {
    if (pe[k] < 0. || dens[k] < 0.) rk[k] = 0., ek[k] = 0., mu[k] = 0., fr[k] = 0.;
  }
  if (e1 != e1 || rho1 != rho1) raiseErr(f.sc, MHY_ERR_NAN);
  // computeEOS on the remapped state
  double pk[M], ck[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const double g = o.gammap[k];
    if (o.eos[k] == MHY_EOS_PERFECT_GAS) {
      pk[k] = (g - 1.0) * rk[k] * ek[k];
      ck[k] = kMixed<R> ? (double)sqrtf(float(g * (g - 1.0) * ek[k])) : sqrt(g * (g - 1.0) * ek[k]);
    } else if (o.eos[k] == MHY_EOS_VOID) {
      pk[k] = (double)at(f.p_k[cur][k], c);
      ck[k] = (double)at(f.c_k[cur][k], c);
      ek[k] = (double)at(f.e_k[cur][k], c);
      e1 = 0.0;
    } else {
      pk[k] = (double)at(f.p_k[nxt][k], c);
      ck[k] = (double)at(f.c_k[nxt][k], c);
    }
  }
  // computePressionMoyenne
  double p1 = 0.0, c1 = 0.0;
  const bool needStale = !(rho1 > 0.);
  if (needStale) c1 = (double)at(f.c[nxt], c);
#pragma unroll
  for (int k = 0; k < M; ++k) {
    p1 += fr[k] * pk[k];
    if (needStale) c1 = smax(c1, ck[k]);
  }
  if (rho1 > 0.) {
    const double c2 = o.gammap[0] * (o.gammap[0] - 1.0) * e1;
    c1 = kMixed<R> ? (double)sqrtf(float(c2)) : sqrt(c2);
  }
#pragma unroll
  for (int k = 0; k < M; ++k)
    if (o.eos[k] == MHY_EOS_VOID) e1 += mu[k] * ek[k];
  out.rho = (R)rho1;
  out.e = (R)e1;
  out.q = (R)q1;
  out.p = (R)p1;
  out.c = (R)c1;
  out.Mc = Mc;
  out.mixte = mixte;
  out.pure = pure;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    out.fr[k] = fr[k];
    out.mu[k] = (R)mu[k];
    out.Mk[k] = Mk[k];
    out.rk[k] = (R)rk[k];
    out.ek[k] = (R)ek[k];
    out.pk[k] = (R)pk[k];
    out.ck[k] = (R)ck[k];
  }
}

template <class R, int M>
__device__ __forceinline__ void finalStore(const KPT<R>& kp, long c, const FinalOut<R, M>& out) {
  const auto& f = kp.f;
  const int nxt = kp.nxt;
  at(f.rho[nxt], c) = out.rho;
  at(f.e[nxt], c) = out.e;
  at(f.q[nxt], c) = out.q;
  at(f.p[nxt], c) = out.p;
  at(f.c[nxt], c) = out.c;
  at(f.M, c) = out.Mc;
  at(f.mixte, c) = out.mixte;
  at(f.pure, c) = out.pure;
  if constexpr (M == 1) {
    if (kp.o.compact1) {
      // the material fields equal 1, 1 and the mean fields bit for bit in
      // every cell but exceptional ones (a negative density or energy, an
      // empty cell): only those store them
      const bool derived = sameBits(out.fr[0], 1.0) && sameBits(out.mu[0], R(1.0)) && sameBits(out.Mk[0], out.Mc) &&
                           sameBits(out.rk[0], out.rho) && sameBits(out.ek[0], out.e) &&
                           sameBits(out.pk[0], out.p) && sameBits(out.ck[0], out.c);
      at(f.envx, c) = derived ? 0 : 1;
      if (derived) return;
    }
  }
#pragma unroll
  for (int k = 0; k < M; ++k) {
    at(f.f_k[k], c) = out.fr[k];
    at(f.mu_k[k], c) = out.mu[k];
    at(f.M_k[k], c) = out.Mk[k];
    at(f.rho_k[nxt][k], c) = out.rk[k];
    at(f.e_k[nxt][k], c) = out.ek[k];
    at(f.p_k[nxt][k], c) = out.pk[k];
    at(f.c_k[nxt][k], c) = out.ck[k];
  }
}

template <class R, int M>
__device__ __forceinline__ void finalCell(const KPT<R>& kp, long c, const double* U) {
  FinalOut<R, M> out;
  finalCompute<R, M>(kp, c, U, out);
  finalStore<R, M>(kp, c, out);
}

}  // namespace mhy
