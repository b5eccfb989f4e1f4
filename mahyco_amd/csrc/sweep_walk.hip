// Remap sweeps along y with one lane per column (Remap::computeGradPhiFace,
// computeGradPhi, computeUpwindFaceQuantitiesForProjection, computeUremap of
// the y direction; remap/PhaseRemap1.cc, remap/PhaseRemap2.cc).
//
// A wave covers 64 consecutive columns and walks down a band of kWalkRows
// rows, plus two halo rows above and below.  Every row access of the wave
// is a whole 64-cell row segment.  Each iteration loads one new row and
// keeps the two before it in registers: the divided difference at the new
// row's minus face, the slope (or the PenteBorne fluxes) of the row before,
// the face value between the two rows before, and the update of the row
// two back.  Operation order is that of the wave-strip sweeps (and of the
// reference), so FP64 results are bitwise the same.
#include <type_traits>

#include "final_math.hpp"
#include "kernels.hpp"
#include "sweep_math.hpp"
#include "sweep_update.hpp"

namespace mhy {

using namespace sw;

namespace {

constexpr int kWalkRows = 32;  // rows updated per band

// end-of-Lagrange node position (updatePosition, PhaseLagrange.cc:203-208)
template <class R>
__device__ __forceinline__ D2 lagNodeK(const KPT<R>& kp, double dt1, long j, long i) {
  const long p = j * (kp.o.nx + 1) + i;
  const D2 e = eulerNode(kp.o, j, i);
  return {e.x + (double)at(kp.f.ux[kp.nxt], p) * dt1, e.y + (double)at(kp.f.uy[kp.nxt], p) * dt1};
}

// width across the y-sweep faces (HvLagrange, PhaseRemap2.cc): right then
// left face
template <class R>
__device__ __forceinline__ double widthY(const KPT<R>& kp, double dt1, long j, long i) {
  double h = 0.;
  h += 0.5 * d2norm(d2sub(lagNodeK(kp, dt1, j + 1, i + 1), lagNodeK(kp, dt1, j, i + 1)));
  h += 0.5 * d2norm(d2sub(lagNodeK(kp, dt1, j + 1, i), lagNodeK(kp, dt1, j, i)));
  return h;
}

}  // namespace

// b: linear block index; the waves of all bands are numbered column-group
// first, 4 per block.
template <class R, int M, bool PP>
__device__ __forceinline__ void sweepWalkYBody(const KPT<R>& kp, unsigned b) {
  using T = typename std::conditional<kMixed<R> && !PP, float, double>::type;
  constexpr int NC = 3 * M + 1;
  constexpr bool kF32 = sizeof(T) == 4;
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const long nx = o.nx, ny = o.ny;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long ngx = (nx + 63) / 64;
  const long g = (long)b * 4 + wave;
  const long i0 = (g % ngx) * 64 + lane;
  const long a0 = (g / ngx) * kWalkRows;

  double* const* Uin = (kp.sweep == 1) ? f.UL : f.U1;
  const double dtn = f.sc->dt[kp.cur], dt1 = f.sc->dt[kp.nxt];
  const bool needH = PP || o.limiter >= MHY_LIM_MINMODG;
  const bool mixteOpt = PP && o.pp_mixte == 1;
  const bool pb = PP && o.methode == 2;
  const bool formed = o.order > 1;
  const R* uy = f.uy[kp.nxt];

  // rolling state: A = row r-2, B = row r-1
  T phA[NC], phB[NC], dMB[NC], gA[NC], avA[NC], faceA[NC], dualA[M];
#pragma unroll
  for (int q = 0; q < NC; ++q) phA[q] = phB[q] = dMB[q] = gA[q] = avA[q] = faceA[q] = T(0.0);
#pragma unroll
  for (int k = 0; k < M; ++k) dualA[k] = T(0.0);
  double xcA = 0.0, xcB = 0.0, vA = 0.0, vB = 0.0;
  T HA = T(0.0), HB = T(0.0);
  int flA = 0, flB = 0;

  for (int k = 0; k < kWalkRows + 4; ++k) {
    // the column and what derives from it (addresses, Euler coordinates) are
    // formed again for every row from an opaque copy of the index: hoisted
    // out of the loop they would hold registers through every row
    int iop = (int)i0;
    asm volatile("" : "+v"(iop));
    const long i = iop;
    const bool inT = (i < nx);
    // row r of this lane's column: Phi, centroid, width, flags
    auto loadRow = [&](long r, T* ph, double& x, T& h, int& flg) {
      const long cb = r * nx + i;
      double u[NC];
      const int nload = PP ? 3 * M : NC;
#pragma unroll
      for (int q = 0; q < NC; ++q) u[q] = (q < nload) ? at(Uin[q], cb) : 0.0;
      cellPhi<M, PP, T>(u, PP ? 1.0 : at(f.VL, cb), PP ? at(f.qL, cb) : 0.0, ph);
      x = at(f.YcL, cb);
      h = needH ? T(widthY(kp, dt1, r, i)) : T(0.0);
      flg = 0;
      if (mixteOpt) flg = (kp.sweep == 1) ? packFlags(at(f.mixte, cb), at(f.pure, cb)) : packedFlags<M, T>(ph);
    };
    // normal velocity of face r, between rows r-1 and r (PrepareRemap.cc:50-62)
    auto fnv = [&](long r) {
      const long n1 = r * (nx + 1) + i, n2 = n1 + 1;
      return 0.0 + ((0.0 + (double)at(uy, n1)) + (double)at(uy, n2)) * 0.5;
    };
    // Euler length of face r (Init.cc:528-533)
    auto faceLen = [&](long r) { return d2norm(d2sub(eulerNode(o, r, i + 1), eulerNode(o, r, i))); };
    const long r = a0 - 2 + k, rB = r - 1, rA = r - 2;
    // new row N
    T phN[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) phN[q] = T(0.0);
    double xcN = 0.0;
    T HN = T(0.0);
    int flN = 0;
    if (inT && r >= 0 && r < ny) loadRow(r, phN, xcN, HN, flN);
    const double vN = (inT && r >= 0 && r <= ny) ? fnv(r) : 0.0;

    // divided difference at the minus face of row N (between rows B and N)
    T dMN[NC];
    {
      const bool inMN = (r >= 1);
      const T xcmT = kF32 ? T(xcB - xcN) : T(xcB), xcT = kF32 ? T(0.0) : T(xcN);
      const Nbr<T> nN{phB, phN, phN, xcmT, xcT, xcT, inMN, false, HN, HB, HN, false};
#pragma unroll
      for (int q = 0; q < NC; ++q) dMN[q] = formed ? faceDiffM<false, T>(nN, q) : T(0.0);
    }

    // row B with its neighbours A and N; a missing neighbour takes row B's
    // values (PhaseRemap1.cc:157-158)
    const bool inB = inT && rB >= 0 && rB < ny;
    const bool inMB = (rB >= 1), inPB = (rB + 1 < ny);
    T phmB[NC], phpB[NC], dPB[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      phmB[q] = inMB ? phA[q] : phB[q];
      phpB[q] = inPB ? phN[q] : phB[q];
      dPB[q] = inPB ? dMN[q] : T(0.0);
    }
    const double xcmB = inMB ? xcA : xcB, xcpB = inPB ? xcN : xcB;
    const T HmB = inMB ? HA : HB, HpB = inPB ? HN : HB;
    const int fmB = inMB ? flA : flB, fpB = inPB ? flN : flB;
    const T xcmT = kF32 ? T(xcmB - xcB) : T(xcmB), xcT = kF32 ? T(0.0) : T(xcB);
    const T xcpT = kF32 ? T(xcpB - xcB) : T(xcpB);
    Nbr<T> nB{phmB, phB, phpB, xcmT, xcT, xcpT, inMB, inPB, HB, HmB, HpB, false};
    if (mixteOpt) nB.vp = ((flB >> 8) == 0) && flB == fpB && flB == fmB;
    nB.dM = dMB;
    nB.dP = dPB;

    // face value at the minus face of row B (between rows A and B)
    T faceB[NC], gB[NC], avB[NC], dualB[M];
#pragma unroll
    for (int q = 0; q < NC; ++q) faceB[q] = gB[q] = avB[q] = T(0.0);
#pragma unroll
    for (int k2 = 0; k2 < M; ++k2) dualB[k2] = T(0.0);
    if (!PP) {
#pragma unroll
      for (int q = 0; q < NC; ++q) gB[q] = inB ? gradQ<false, T>(o, nB, q) : T(0.0);
      // upwind value (UtilesRemap-Impl.h:491-501); y: back = above (row B),
      // front = below (row A)
      const double dxl = xcA - xcB;  // front - back
      const bool upLo = !(vB * dxl > 0.0);
      const double xf = 0.5 * (o.dy * (double)rB + o.dy * (double)rB);
      const T s = T(0.0 + (xf - (upLo ? xcA : xcB)));
#pragma unroll
      for (int q = 0; q < NC; ++q) faceB[q] = upLo ? phA[q] + gA[q] * s : phB[q] + gB[q] * s;
    } else {
      T arB[NC];
#pragma unroll
      for (int q = 0; q < NC; ++q) arB[q] = T(0.0);
      if (formed && inB) ppFluxes<M, false, T>(o, nB, T(vB), T(vN), T(dtn), pb, arB, avB, dualB);
#pragma unroll
      for (int q = 0; q < NC; ++q) faceB[q] = formed ? avA[q] - arB[q] : T(0.0);
    }

    // update of row A: minus face faceA, plus face faceB
    if (inT && rA >= a0 && rA < a0 + kWalkRows && rA < ny) {
      const long c = rA * nx + i;
      const bool hasM = (rA >= 1), hasP = (rA + 1 < ny);
      const double Lm = faceLen(rA), Lp = faceLen(rA + 1);
      T phiM[NC], phiP[NC];
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        phiM[q] = hasM ? faceA[q] : T(0.0);
        phiP[q] = hasP ? faceB[q] : T(0.0);
      }
      if (pb) {
        const double w = (Lm + Lp) * 0.5;
#pragma unroll
        for (int k2 = 0; k2 < M; ++k2) at(f.DF[k2], c) = formed ? (double)dualA[k2] * w : 0.0;
      }
      sweepUpdateCell<R, M, PP, false, T>(kp, c, i, rA, rA, hasM, hasP, vA, vB, Lm, Lp, phiM, phiP);
    }

    // rotate: B becomes A, N becomes B
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      phA[q] = phB[q], phB[q] = phN[q];
      dMB[q] = dMN[q];
      gA[q] = gB[q];
      avA[q] = avB[q];
      faceA[q] = faceB[q];
    }
#pragma unroll
    for (int k2 = 0; k2 < M; ++k2) dualA[k2] = dualB[k2];
    xcA = xcB, xcB = xcN;
    HA = HB, HB = HN;
    flA = flB, flB = flN;
    vA = vB, vB = vN;
  }
}

// waves per SIMD the walker is compiled for: 4 for one material with limited
// slopes (128 VGPRs, a few spilled; unbounded it takes 136 and runs 3 waves)
template <class R, int M, bool PP>
constexpr int kWalkWaves = (M == 1 && !PP) ? 4 : 1;

template <class R, int M, bool PP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kWalkWaves<R, M, PP>))) void k_sweep_walk(
    KPT<R> kp) {
  sweepWalkYBody<R, M, PP>(kp, blockIdx.x);
}

template <class R, int M, bool PP, bool ONE>
__global__ __launch_bounds__(256) void k_sweep_walk_batch(BatchLaunchT<R> bl) {
  batchRun<ONE>(bl, [&](const KPT<R>& kp, unsigned local, unsigned, unsigned) { sweepWalkYBody<R, M, PP>(kp, local); });
}

#define MHY_SWEEP_WALK_INST(R, M, PP)                                       \
  template __global__ void k_sweep_walk<R, M, PP>(KPT<R>);                  \
  template __global__ void k_sweep_walk_batch<R, M, PP, false>(BatchLaunchT<R>); \
  template __global__ void k_sweep_walk_batch<R, M, PP, true>(BatchLaunchT<R>);
#define MHY_SWEEP_WALK_INST_R(R)    \
  MHY_SWEEP_WALK_INST(R, 1, false) \
  MHY_SWEEP_WALK_INST(R, 1, true)  \
  MHY_SWEEP_WALK_INST(R, 2, false) \
  MHY_SWEEP_WALK_INST(R, 2, true)  \
  MHY_SWEEP_WALK_INST(R, 3, false) \
  MHY_SWEEP_WALK_INST(R, 3, true)
MHY_SWEEP_WALK_INST_R(double)
MHY_SWEEP_WALK_INST_R(float)
#undef MHY_SWEEP_WALK_INST_R
#undef MHY_SWEEP_WALK_INST

// launch grid of the walking y sweeps (256-thread blocks of 4 waves)
dim3 sweepWalkGrid(long nx, long ny) {
  const long waves = (nx + 63) / 64 * ((ny + kWalkRows - 1) / kWalkRows);
  return dim3((unsigned)((waves + 3) / 4));
}

}  // namespace mhy
