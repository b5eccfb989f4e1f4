// Update of one cell by the two face fluxes of a remap sweep
// (computeUremap{1,2}, PhaseRemap1.cc:459-568, PhaseRemap2.cc:454-497, with
// the boundary fluxes of remap/BoundaryFluxes.cc:14-120), then either the
// store of the new remap vector or, in the second sweep, the final remap of
// the cell from registers; and the face fluxes the dual remap needs.
// Shared by the wave-strip sweeps and the column-walking y sweeps.
#pragma once

#include <type_traits>

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
// vm / vp and Euler face lengths Lm / Lp: its new remap vector Unew and the
// flux of its minus face for the dual remap (A1: mass, A2: volume; zero for
// PB).  The flux of the last face of a line is stored here.  Uown: the
// cell's remap vector at the start of the sweep when the caller holds it in
// registers (the first sweep with the Lagrange update fused), else null and
// it is read from U_L / U_1.  L: the vector layout (kMat, kNC).
template <class R, int L, bool PP, bool XD, class T>
__device__ __forceinline__ void sweepUnew(const KPT<R>& kp, long c, long i, long j, long a, bool hasM, bool hasP,
                                          double vm, double vp, double Lm, double Lp, const T* phiM, const T* phiP,
                                          double* Unew, double* Gdual, const double* Uown = nullptr) {
  constexpr int NC = kNC<L>;
  constexpr int M = kMat<L>;
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const long nx = o.nx, ny = o.ny;
  double* const* Uin = (kp.sweep == 1) ? f.UL : f.U1;
  auto uOld = [&](int q) { return Uown ? Uown[q] : at(Uin[q], c); };
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
  if (kMixed<R>) {
    // U - (-Gm + Gp + Fb), compensated: the FP64 result is the correctly
    // rounded sum up to a final rounding, whatever the flux magnitudes
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      double sum = uOld(q), comp = 0.0;
      if (hasM) twoSumAcc(sum, comp, Gm[q]);
      if (hasP) twoSumAcc(sum, comp, -Gp[q]);
      twoSumAcc(sum, comp, -Fb[q]);
      Unew[q] = sum + comp;
    }
  } else {
#pragma unroll
    for (int q = 0; q < NC; ++q) Unew[q] = uOld(q) - red[q];
  }
  const bool withG = (o.methode == 0 || o.methode == 1);
  const int q0 = (o.methode == 0) ? M : 0;
#pragma unroll
  for (int k = 0; k < M; ++k) Gdual[k] = (withG && hasM) ? Gm[q0 + k] : 0.0;
  if (withG && !hasP) {
    const long fp = XD ? j * (nx + 1) + a + 1 : (a + 1) * nx + i;
#pragma unroll
    for (int k = 0; k < M; ++k) f.G[k][fp] = 0.0;
  }
}

// The update of cell c and its stores: the final remap of the cell from
// registers (second sweep, fused; VNR) or the new remap vector, and the flux
// of its minus face for the dual remap.
template <class R, int L, bool PP, bool XD, class T>
__device__ __forceinline__ void sweepUpdateCell(const KPT<R>& kp, long c, long i, long j, long a, bool hasM,
                                                bool hasP, double vm, double vp, double Lm, double Lp,
                                                const T* phiM, const T* phiP, const double* Uown = nullptr) {
  constexpr int NC = kNC<L>;
  constexpr int M = kMat<L>;
  const auto& f = kp.f;
  double Unew[NC], Gd[M];
  sweepUnew<R, L, PP, XD, T>(kp, c, i, j, a, hasM, hasP, vm, vp, Lm, Lp, phiM, phiP, Unew, Gd, Uown);
  bool stored = false;
  if constexpr (!kEuc<L>) {
    if (kp.fuseFinal) {
      finalCell<R, M>(kp, c, Unew);  // remapVariables on the cell, U2 stays in registers
      stored = true;
    }
  }
  if (!stored) {
    double* const* Uout = (kp.sweep == 1) ? f.U1 : f.U2;
#pragma unroll
    for (int q = 0; q < NC; ++q) at(Uout[q], c) = Unew[q];
  }
  if (kp.o.methode == 0 || kp.o.methode == 1) {
    const long fm = XD ? j * (kp.o.nx + 1) + a : a * kp.o.nx + i;
#pragma unroll
    for (int k = 0; k < M; ++k) at(f.G[k], fm) = Gd[k];
  }
}

// What a sweep stores for one cell: the final remap's fields (FUSE) or the
// new remap vector, the minus-face flux for the dual remap and the PB dual
// flux.  The sweep order is a template parameter wherever a CellOut is
// used: a run-time choice between its members makes the compiler merge the
// two paths' accesses through a selected address, which puts the struct in
// scratch memory.
template <class R, int L>
struct CellOut {
  FinalOut<R, kMat<L>> fin;
  double U[kNC<L>];
  double G[kMat<L>];
  double DF[kMat<L>];
};

template <class R, int L, bool FUSE>
__device__ __forceinline__ void storeCellOut(const KPT<R>& kp, long c, long fm, const CellOut<R, L>& out, bool withG) {
  const auto& f = kp.f;
  if constexpr (FUSE) {
    finalStore<R, kMat<L>>(kp, c, out.fin);
  } else {
    double* const* Uout = (kp.sweep == 1) ? f.U1 : f.U2;
#pragma unroll
    for (int q = 0; q < kNC<L>; ++q) at(Uout[q], c) = out.U[q];
  }
  if (withG) {
#pragma unroll
    for (int k = 0; k < kMat<L>; ++k) at(f.G[k], fm) = out.G[k];
  }
}

// Visits the members of out that a launch stores, in a fixed order: the
// final fields (FUSE) or the remap vector, then G (withG), then DF (withDF).
template <class R, int L, bool FUSE, class Fn>
__device__ __forceinline__ void visitCellOut(CellOut<R, L>& out, bool withG, bool withDF, Fn&& fn) {
  constexpr int M = kMat<L>;
  if constexpr (FUSE) {
    FinalOut<R, M>& fo = out.fin;
    fn(fo.rho), fn(fo.e), fn(fo.q), fn(fo.p), fn(fo.c), fn(fo.mixte), fn(fo.pure), fn(fo.Mc);
#pragma unroll
    for (int k = 0; k < M; ++k) fn(fo.fr[k]), fn(fo.Mk[k]), fn(fo.mu[k]), fn(fo.rk[k]), fn(fo.ek[k]), fn(fo.pk[k]), fn(fo.ck[k]);
  } else {
#pragma unroll
    for (int q = 0; q < kNC<L>; ++q) fn(out.U[q]);
  }
  if (withG) {
#pragma unroll
    for (int k = 0; k < M; ++k) fn(out.G[k]);
  }
  if (withDF) {
#pragma unroll
    for (int k = 0; k < M; ++k) fn(out.DF[k]);
  }
}

// 8- and 4-byte members of the final fields
template <class R, int M>
constexpr int kFuse8 = 1 + 2 * M + (sizeof(R) == 8 ? 5 + 5 * M : 0);
template <class R, int M>
constexpr int kFuse4 = 2 + (sizeof(R) == 4 ? 5 + 5 * M : 0);
// LDS bytes per staged cell: the larger of the two layouts (final fields or
// remap vector), each followed by G and DF
template <class R, int L, bool PP>
constexpr int kStageFuseBytes = (kFuse8<R, kMat<L>> + kMat<L> + (PP ? kMat<L> : 0)) * 8 + kFuse4<R, kMat<L>> * 4;
template <class R, int L, bool PP>
constexpr int kStageVecBytes = (kNC<L> + kMat<L> + (PP ? kMat<L> : 0)) * 8;
// (the Eucclhyd layouts never stage the final fields)
template <class R, int L, bool PP>
constexpr int kStageBytes = (!kEuc<L> && kStageFuseBytes<R, L, PP> > kStageVecBytes<R, L, PP>)
                                ? kStageFuseBytes<R, L, PP>
                                : kStageVecBytes<R, L, PP>;
template <class R, int L, bool PP>
constexpr int kStageWords = (kStageBytes<R, L, PP> + 7) / 8;

// LDS staging of the CellOut of a tile of TILE cells, in a buffer of
// kStageWords * TILE doubles: 8-byte member w of element e at
// s8[w * TILE + e], then the 4-byte members at s4[w * TILE + e]
template <int TILE>
struct OutStage {
  double* s8;
  unsigned* s4;
  template <class R, int L, bool FUSE>
  __device__ __forceinline__ static OutStage make(double* buf, bool withG, bool withDF) {
    constexpr int M = kMat<L>;
    const int n8 = (FUSE ? kFuse8<R, M> : kNC<L>) + (withG ? M : 0) + (withDF ? M : 0);
    return OutStage{buf, reinterpret_cast<unsigned*>(buf + n8 * TILE)};
  }
  template <class R, int L, bool FUSE>
  __device__ __forceinline__ void put(CellOut<R, L>& out, int e, bool withG, bool withDF) const {
    int w8 = 0, w4 = 0;
    visitCellOut<R, L, FUSE>(out, withG, withDF, [&](auto& v) {
      using V = typename std::remove_reference<decltype(v)>::type;
      if constexpr (sizeof(V) == 8) s8[(w8++) * TILE + e] = __builtin_bit_cast(double, v);
      else s4[(w4++) * TILE + e] = __builtin_bit_cast(unsigned, v);
    });
  }
  template <class R, int L, bool FUSE>
  __device__ __forceinline__ void get(CellOut<R, L>& out, int e, bool withG, bool withDF) const {
    int w8 = 0, w4 = 0;
    visitCellOut<R, L, FUSE>(out, withG, withDF, [&](auto& v) {
      using V = typename std::remove_reference<decltype(v)>::type;
      if constexpr (sizeof(V) == 8) v = __builtin_bit_cast(V, s8[(w8++) * TILE + e]);
      else v = __builtin_bit_cast(V, s4[(w4++) * TILE + e]);
    });
  }
};

}  // namespace mhy
