// Update of one cell by the two face fluxes of a remap sweep
// (computeUremap{1,2}, PhaseRemap1.cc:459-568, PhaseRemap2.cc:454-497, with
// the boundary fluxes of remap/BoundaryFluxes.cc:14-120), then either the
// store of the new remap vector or, in the second sweep, the final remap of
// the cell from registers; and the face fluxes the dual remap needs.
// Shared by the wave-strip sweeps and the column-walking y sweeps.
#pragma once

#include "common.hpp"
#include "final_math.hpp"

namespace mhy {

// s + v with the rounding error of the addition added to comp (TwoSum)
__device__ __forceinline__ void twoSumAcc(double& s, double& comp, double v) {
  const double t = s + v;
  const double bv = t - s;
  comp += (s - (t - bv)) + (v - bv);
  s = t;
}

// Cell c (along index a, row j, column i) with face values phiM / phiP at
// its minus / plus faces (valid when hasM / hasP), face normal velocities
// vm / vp and Euler face lengths Lm / Lp.
template <class R, int M, bool PP, bool XD, class T>
__device__ __forceinline__ void sweepUpdateCell(const KPT<R>& kp, long c, long i, long j, long a, bool hasM,
                                                bool hasP, double vm, double vp, double Lm, double Lp,
                                                const T* phiM, const T* phiP) {
  constexpr int NC = 3 * M + 1;
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const long nx = o.nx, ny = o.ny;
  double* const* Uin = (kp.sweep == 1) ? f.UL : f.U1;
  double* const* Uout = (kp.sweep == 1) ? f.U1 : f.U2;
  const double dtn = f.sc->dt[kp.cur];
  // computeUremap (PhaseRemap1.cc:468-503): a cell receives -G through its
  // minus face and +G through its plus face
  const double coefM = PP ? Lm : vm * Lm * dtn;
  const double coefP = PP ? Lp : vp * Lp * dtn;
  double Gm[NC], Gp[NC], red[NC], Fb[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    Gm[q] = coefM * (double)phiM[q];
    Gp[q] = coefP * (double)phiP[q];
    red[q] = 0.0;
    Fb[q] = 0.0;
    if (hasM) red[q] = red[q] + (-Gm[q]);
    if (hasP) red[q] = red[q] + Gp[q];
  }
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
      const bool oppOk = viaMinus ? hasP : hasM;
      const double vb = viaMinus ? vm : vp;
      const double Lb = viaMinus ? Lm : Lp;
      const double sgn = viaMinus ? -1.0 : 1.0;
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        // values, not a conditional lvalue: a select between the two arrays'
        // addresses would keep them in scratch memory
        const double pP = phiP[q], pM = phiM[q];
        const double ph = oppOk ? (viaMinus ? pP : pM) : 0.0;
        Fb[q] = PP ? (sgn * Lb) * ph : (((sgn * vb) * Lb) * dtn) * ph;
        red[q] = red[q] + Fb[q];
      }
    }
  }
  double Unew[NC];
  if (kMixed<R>) {
    // U - (-Gm + Gp + Fb), compensated: the FP64 result is the correctly
    // rounded sum up to a final rounding, whatever the flux magnitudes
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      double sum = Uin[q][c], comp = 0.0;
      if (hasM) twoSumAcc(sum, comp, Gm[q]);
      if (hasP) twoSumAcc(sum, comp, -Gp[q]);
      twoSumAcc(sum, comp, -Fb[q]);
      Unew[q] = sum + comp;
    }
  } else {
#pragma unroll
    for (int q = 0; q < NC; ++q) Unew[q] = Uin[q][c] - red[q];
  }
  if (kp.fuseFinal) {
    finalCell<R, M>(kp, c, Unew);  // remapVariables on the cell, U2 stays in registers
  } else {
#pragma unroll
    for (int q = 0; q < NC; ++q) Uout[q][c] = Unew[q];
  }

  // face fluxes for the dual remap (A1: mass, A2: volume): the minus face of
  // each cell, plus the last face of the line
  if (o.methode == 0 || o.methode == 1) {
    const int q0 = (o.methode == 0) ? M : 0;
    const long fm = XD ? j * (nx + 1) + a : a * nx + i;
#pragma unroll
    for (int k = 0; k < M; ++k) f.G[k][fm] = hasM ? Gm[q0 + k] : 0.0;
    if (!hasP) {
      const long fp = XD ? j * (nx + 1) + a + 1 : (a + 1) * nx + i;
#pragma unroll
      for (int k = 0; k < M; ++k) f.G[k][fp] = 0.0;
    }
  }
}

}  // namespace mhy
