// Cell part of the VNR Lagrange phase for one cell: updateCellPos,
// computeSubVol, updateRho (+ periodic cells), computeTau, computeDivU,
// computeArtificialViscosity, updateEnergy, computeEOS,
// computePressionMoyenne (vnr_remap/PhaseLagrange.cc:214-485,
// vnr_remap/ConditionsLimites.cc:186-209), then the cell part of the remap
// preparation: computeVariablesForRemap and computeCellQuantitesForRemap
// (vnr_remap/PrepareRemap.cc:65-176).
//
// Shared by the stand-alone cell kernel (lagrange.hip), which stores the
// remap vector for the first sweep, and by the first remap sweep, which
// forms it in registers instead (sweep_wave.hip).  The Lagrange volume and
// centroid depend only on the end-of-Lagrange node positions, so the second
// sweep forms them again from the node velocities (lagGeometry) rather than
// reading them back.  Arithmetic is FP64 in both forms, in the reference's
// operation order (compiled with -ffp-contract=off).
#pragma once

#include "common.hpp"
#include "final_math.hpp"

namespace mhy {

// End-of-Lagrange positions of the nodes of cell (j, i), order BL, BR, TR,
// TL: the moving-mesh coordinates, or the Euler node moved by dt1 * u^{n+1}
// (updatePosition, PhaseLagrange.cc:203-208).
template <class R>
__device__ __forceinline__ void lagNodes(const KPT<R>& kp, double dt1, long j, long i, D2* X) {
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const long nx = o.nx;
  const long jj[4] = {j, j, j + 1, j + 1}, ii[4] = {i, i + 1, i + 1, i};
  // the corners' node indices differ from the first by uniform amounts,
  // which go into the array bases: one 32-bit offset serves the four loads
  const long p0 = j * (nx + 1) + i;
  const long dp[4] = {0, 1, nx + 2, nx + 1};
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (o.mode == MODE_LAGRANGE) {
      X[s] = {at(f.X[kp.nxt] + dp[s], p0), at(f.Y[kp.nxt] + dp[s], p0)};
    } else {
      const D2 e = eulerNode(o, jj[s], ii[s]);
      X[s] = {e.x + (double)at(f.ux[kp.nxt] + dp[s], p0) * dt1, e.y + (double)at(f.uy[kp.nxt] + dp[s], p0) * dt1};
    }
  }
}

// Cell volume as the sum of its corner sub-volumes around the node mean
// (updateCellPos + computeSubVol + the sum of updateRho).
__device__ __forceinline__ double lagVolume(const D2* X) {
  D2 xc = {0.0, 0.0};
#pragma unroll
  for (int s = 0; s < 4; ++s) xc = d2add(xc, X[s]);
  xc = d2mul(xc, 0.25);
  double V = 0.0;
#pragma unroll
  for (int s = 0; s < 4; ++s) V = V + subVolume(xc, X[(s + 3) & 3], X[s], X[(s + 1) & 3]);
  return V;
}

// Polygon centroid of volume V (computeCellQuantitesForRemap, PrepareRemap.cc:65-90)
__device__ __forceinline__ D2 lagCentroid(const D2* X, double V) {
  D2 r = {0.0, 0.0};
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const D2 a = X[s], b = X[(s + 1) & 3];
    r = d2add(r, d2mul(d2add(a, b), d2cross(a, b)));
  }
  return d2mul(r, 1.0 / (6.0 * V));
}

// Lagrange volume and centroid of cell (j, i), as the cell update forms them.
template <class R>
__device__ __forceinline__ void lagGeometry(const KPT<R>& kp, double dt1, long j, long i, double& V, D2& xc) {
  D2 X[4];
  lagNodes(kp, dt1, j, i, X);
  V = lagVolume(X);
  xc = lagCentroid(X, V);
}

// What the remap needs from the Lagrange update of one cell: the live
// components of the remap vector U_L, the Lagrange volume and centroid, and
// the Lagrange pseudo-viscosity and density (PenteBorne Phi, A2 dual fluxes).
template <int M>
struct LagCell {
  double U[3 * M + 1];
  double V, q, rho;
  D2 xc;
};

// The cell update of cell c = (j, i).  store: write the cell's n+1 words
// that outlive the step (div u, the per-material pseudo; in the moving-mesh
// mode every Lagrange output).  Returns false in the moving-mesh mode, where
// nothing is remapped and out is left unset.  The mixed form takes the remap
// masses from the FP64 material masses M_k (equal to mu_k V rho in exact
// arithmetic), so that mass stays conserved to FP64 round-off whatever the
// FP32 density storage (periodic cells, whose density is reset to its
// initial value, keep mu_k V rho).
template <class R, int M>
__device__ __forceinline__ bool lagrangeCellCompute(const KPT<R>& kp, long c, long j, long i, bool store,
                                                    LagCell<M>& out) {
  const Opts& o = kp.o;
  const auto& f = kp.f;
  constexpr int m = M;  // material count (template: the loops unroll)
  const long nx = o.nx;
  const int cur = kp.cur, nxt = kp.nxt;
  const Scalars* sc = f.sc;
  const double dt1 = sc->dt[nxt];
  // the cell's step-start words, loaded up front: their uses below are
  // conditional, and unconditional loads can all be in flight together
  double fk[NMAT], Mk0[NMAT], rk0[NMAT], qk0[NMAT], ek0[NMAT], pk0[NMAT], mu0[NMAT];
  const double rhoS = (double)at(f.rho[cur], c), divuS = (double)at(f.divu[cur], c), cS = (double)at(f.c[cur], c);
  const double vol0 = at(f.vol[cur], c), Mc0 = at(f.M, c);
  // one material (compact1): fraction and mass fraction 1, the material
  // mass, density, energy and pressure those of the cell, unless the final
  // remap stored them (envx)
  const bool derived = (m == 1) && o.compact1 && !at(f.envx, c);
  if (derived) {
    fk[0] = 1.0, mu0[0] = 1.0, Mk0[0] = Mc0, rk0[0] = rhoS;
    ek0[0] = (double)at(f.e[cur], c), pk0[0] = (double)at(f.p[cur], c);
    qk0[0] = (double)at(f.q_k[cur][0], c);
  } else {
#pragma unroll
    for (int k = 0; k < m; ++k) {
      fk[k] = at(f.f_k[k], c);
      Mk0[k] = at(f.M_k[k], c);
      rk0[k] = (double)at(f.rho_k[cur][k], c);
      qk0[k] = (double)at(f.q_k[cur][k], c);
      ek0[k] = (double)at(f.e_k[cur][k], c);
      pk0[k] = (double)at(f.p_k[cur][k], c);
      mu0[k] = (double)at(f.mu_k[k], c);
    }
  }
  D2 X[4];
  lagNodes(kp, dt1, j, i, X);
  const double V = lagVolume(X);

  double rk1[NMAT];
  double rho1 = 0.0;
#pragma unroll
  for (int k = 0; k < m; ++k) {
    // absent materials keep the stale buffer value in the reference; their
    // volume fraction is exactly zero, so the value never contributes
    rk1[k] = (fk[k] > o.threshold) ? Mk0[k] / (fk[k] * V)
                                   : (o.mode == MODE_FULL ? 0.0 : (double)at(f.rho_k[nxt][k], c));
    rho1 += fk[k] * rk1[k];
  }
  const bool periodicCell = o.periodic_right && (i == 0 || i == nx - 1);
  if (periodicCell) {
    const long L = j * nx;  // updatePeriodicBoundaryConditions
    rho1 = (double)at(f.rho0, L);
#pragma unroll
    for (int k = 0; k < m; ++k) rk1[k] = at(f.rho_k0[k], L);
  }

  double q1 = 0.0, e1 = 0.0;
  if (o.mode != MODE_ADVECTION) {
    // mixed form: the step-start densities from the FP64 masses and volume,
    // so that 1/rho1 - 1/rho0 (divided by dt in div u) carries no FP32
    // rounding of the stored density
    const double rho0 = kMixed<R> ? Mc0 / vol0 : rhoS;
    const double tau1 = 0.5 * (1.0 / rho1 + 1.0 / rho0);
    const double div1 = 1.0 / dt1 * (1.0 / rho1 - 1.0 / rho0) / tau1;
    if (divuS < 0.0) {
      const double r0 = vol0;
      q1 = 1.0 / tau1 * (-0.5 * sqrt(r0) * cS * div1 + (o.gamma + 1) / 2.0 * r0 * div1 * div1);
    } else {
      q1 = 0.0;
    }
    const double P = sc->pmax;
    if (q1 > P) q1 = P;
    double qk1[NMAT], ek1[NMAT], pk1[NMAT], ck1[NMAT];
#pragma unroll
    for (int k = 0; k < m; ++k) qk1[k] = fk[k] * q1;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < m; ++k) {
      ek1[k] = 0.0;
      const double r1 = rk1[k];
      const double r0 = (kMixed<R> && fk[k] > o.threshold) ? Mk0[k] / (fk[k] * vol0) : rk0[k];
      if ((r1 > o.threshold) && (r0 > o.threshold)) {
        const double qn = qk0[k];
        double pseudo = 0.0;
        if ((o.pseudo_centree == 1) && ((qk1[k] + qn) * (1.0 / r1 - 1.0 / r0) > 0.)) pseudo = 0.5 * (qk1[k] + qn);
        if (o.pseudo_centree == 0) pseudo = qk1[k];
        const double den = 1 + 0.5 * (o.gammap[k] - 1.0) * r1 * (1.0 / r1 - 1.0 / r0);
        const double num = ek0[k] - (0.5 * pk0[k] + pseudo) * (1.0 / r1 - 1.0 / r0);
        ek1[k] = num / den;
        e1 += mu0[k] * ek1[k];
        if (fabs(ek1[k]) > 5. && o.abort_energy) bad = true;
      }
    }
    if (bad) raiseErr(f.sc, MHY_ERR_ENERGY);
    // computeEOS: perfect gas and void (PhaseLagrange.cc:392-430)
#pragma unroll
    for (int k = 0; k < m; ++k) {
      const double g = o.gammap[k];
      if (o.eos[k] == MHY_EOS_PERFECT_GAS) {
        pk1[k] = (g - 1.0) * rk1[k] * ek1[k];
        ck1[k] = sqrt(g * (g - 1.0) * ek1[k]);
      } else if (o.eos[k] == MHY_EOS_VOID) {
        pk1[k] = pk0[k];
        ck1[k] = at(f.c_k[cur][k], c);
        ek1[k] = ek0[k];
        e1 = 0.0;
      } else {
        pk1[k] = at(f.p_k[nxt][k], c);
        ck1[k] = at(f.c_k[nxt][k], c);
      }
    }
    // computePressionMoyenne (PhaseLagrange.cc:464-485)
    double p1 = 0.0, c1 = 0.0;
    const bool needStale = !(rho1 > 0.);
    if (needStale) c1 = (double)at(f.c[nxt], c);
#pragma unroll
    for (int k = 0; k < m; ++k) {
      p1 += fk[k] * pk1[k];
      if (needStale) c1 = smax(c1, ck1[k]);
    }
    if (rho1 > 0.) c1 = sqrt(o.gammap[0] * (o.gammap[0] - 1.0) * e1);
#pragma unroll
    for (int k = 0; k < m; ++k)
      if (o.eos[k] == MHY_EOS_VOID) e1 += mu0[k] * ek1[k];
    if (store) {
      at(f.divu[nxt], c) = div1;
#pragma unroll
      for (int k = 0; k < m; ++k) at(f.q_k[nxt][k], c) = qk1[k];
    }
    if (o.mode == MODE_LAGRANGE) {
      if (store) {
        at(f.rho[nxt], c) = rho1;
        at(f.q[nxt], c) = q1;
        at(f.e[nxt], c) = e1;
        at(f.p[nxt], c) = p1;
        at(f.c[nxt], c) = c1;
#pragma unroll
        for (int k = 0; k < m; ++k) {
          at(f.rho_k[nxt][k], c) = rk1[k];
          at(f.e_k[nxt][k], c) = ek1[k];
          at(f.p_k[nxt][k], c) = pk1[k];
          at(f.c_k[nxt][k], c) = ck1[k];
        }
        at(f.vol[nxt], c) = V;  // node_cellvolume_n swap (VnrRemap.cc:217-221)
      }
      return false;
    }
  } else {
    // advection: the energy swap of VnrRemap.cc:170 makes the remap read the
    // current energy; the pseudo buffer of the n+1 slot is read as left by
    // the previous remap of that slot
    e1 = (double)at(f.e[cur], c);
    q1 = (double)at(f.q[nxt], c);
  }
  // computeVariablesForRemap (PrepareRemap.cc:96-176), live components only
#pragma unroll
  for (int k = 0; k < m; ++k) {
    out.U[k] = fk[k] * V;
    if (kMixed<R> && !periodicCell) {
      const double Mk = Mk0[k];
      out.U[m + k] = Mk;
      out.U[2 * m + k] = Mk * e1;
    } else {
      const double mu = mu0[k];
      out.U[m + k] = mu * V * rho1;
      out.U[2 * m + k] = mu * V * rho1 * e1;
    }
  }
  out.U[3 * m] = q1 * V;
  out.V = V;
  out.q = q1;
  out.rho = rho1;
  out.xc = lagCentroid(X, V);
  return true;
}

}  // namespace mhy
