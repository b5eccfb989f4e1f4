// Remap sweeps (Remap::computeGradPhiFace{1,2}, computeGradPhi{1,2},
// computeUpwindFaceQuantitiesForProjection{1,2}, computeUremap{1,2};
// remap/PhaseRemap1.cc, remap/PhaseRemap2.cc), one lane per cell, with the
// neighbours' values exchanged inside the 64-wide wavefront.
//
// A wave covers a strip of cells along the sweep direction plus two halo
// cells at each end: along x, 64 consecutive cells of a row (60 updated);
// along y, 4 columns x 16 rows (12 rows updated), so that every row access
// still reads 4 consecutive cells and a 256-thread block reads 16.  Each lane
// loads only its own cell, forms the divided difference at its minus face,
// its own limited slope (or its own PenteBorne fluxes Ar and Av) once, and
// takes the neighbour values it needs with wave shuffles: no LDS is
// allocated, and every slope an updated cell needs is formed inside the wave
// (no lane loads a neighbour from memory on its own).  Operation order follows the reference, so
// FP64 results are bitwise those of the CPU path.
//
// R is the bulk storage type.  With R = float (mixed form) the limited-slope
// sweeps compute Phi, slopes and face values in FP32, with centroid offsets
// formed in FP64 before rounding (so the FP32 slopes see cell-size
// differences, not absolute coordinates); the PenteBorne sweeps keep FP64
// arithmetic, because their swept integrals of nearly emptied materials
// cancel catastrophically in FP32.  Each face flux is formed in FP64 and
// accumulated into the FP64 remap vector with a compensated (TwoSum) update.
//
// LAG: the first sweep of a step also runs the cell part of the Lagrange
// phase (lagrangeCellCompute): each lane forms its cell's Lagrange update and
// remap vector in registers, the halo lanes redundantly, and only the
// updated cells store the words that outlive the step.  The remap vector,
// the Lagrange volume and the centroid never go through memory; the second
// sweep forms the volume and centroid again from the node velocities.
#include <type_traits>

#include "final_math.hpp"
#include "kernels.hpp"
#include "lagrange_cell.hpp"
#include "sweep_math.hpp"
#include "sweep_update.hpp"

namespace mhy {

using namespace sw;

namespace {

constexpr int kWave = 64;
constexpr int kHalo = 2;                // halo cells at each end of a strip
constexpr int kXOut = kWave - 2 * kHalo;   // cells updated per wave along x
constexpr int kYCols = 4;               // columns per wave along y
// waves per block along y: 16 columns, or 32 in the mixed form's
// limited-slope sweeps (FP32 rows, few registers), so that a block writes
// whole 128-byte lines of each row
template <class R, bool PP>
constexpr int kYWaves = (kMixed<R> && !PP) ? 8 : 4;
constexpr int kYRows = kWave / kYCols;
constexpr int kYOut = kYRows - 2 * kHalo;  // rows updated per wave along y
// the fused Lagrange sweeps that park the remap vector in LDS and are
// compiled for 5 waves per SIMD: one material, FP64, limited slopes
template <class R, int L, bool PP>
constexpr bool kParkLag = !kMixed<R> && L == 1 && !PP;

// end-of-Lagrange node position (updatePosition, PhaseLagrange.cc:203-208)
template <class R>
__device__ __forceinline__ D2 lagNodeW(const KPT<R>& kp, double dt1, long j, long i) {
  const long p = j * (kp.o.nx + 1) + i;
  const D2 e = eulerNode(kp.o, j, i);
  return {e.x + (double)at(kp.f.ux[kp.nxt], p) * dt1, e.y + (double)at(kp.f.uy[kp.nxt], p) * dt1};
}

// width across the sweep faces (HvLagrange, PhaseRemap1.cc:57-71, :104-119)
template <bool XD, class R>
__device__ __forceinline__ double widthW(const KPT<R>& kp, double dt1, long j, long i) {
  double h = 0.;
  if (XD) {  // bottom then top face
    h += 0.5 * d2norm(d2sub(lagNodeW(kp, dt1, j, i + 1), lagNodeW(kp, dt1, j, i)));
    h += 0.5 * d2norm(d2sub(lagNodeW(kp, dt1, j + 1, i + 1), lagNodeW(kp, dt1, j + 1, i)));
  } else {  // right then left face
    h += 0.5 * d2norm(d2sub(lagNodeW(kp, dt1, j + 1, i + 1), lagNodeW(kp, dt1, j, i + 1)));
    h += 0.5 * d2norm(d2sub(lagNodeW(kp, dt1, j + 1, i), lagNodeW(kp, dt1, j, i)));
  }
  return h;
}

__device__ __forceinline__ double shfl(double v, int src) { return __shfl(v, src, kWave); }
__device__ __forceinline__ float shfl(float v, int src) { return __shfl(v, src, kWave); }
__device__ __forceinline__ int shfl(int v, int src) { return __shfl(v, src, kWave); }

// value of the along-sweep minus / plus neighbour lane (lm / lp: its index,
// or the lane itself at the wave's ends).  Along x the neighbours are the
// adjacent lanes (DPP); along y they are kYCols lanes away (shuffle).
// Called in converged control flow only.
template <bool XD, class V>
__device__ __forceinline__ V nbrM(V v, int lm) {
  if constexpr (XD) return dppMove<kDppWaveShr1>(v);
  else return shfl(v, lm);
}
template <bool XD, class V>
__device__ __forceinline__ V nbrP(V v, int lp) {
  if constexpr (XD) return dppMove<kDppWaveShl1>(v);
  else return shfl(v, lp);
}

}  // namespace

// Along y a lane's row segment is only 4 cells wide, so the block stages the
// outputs of its 12 x 4W tile in LDS and then stores them with each thread on
// one cell, consecutive threads on consecutive cells of a row: every store
// instruction writes whole 128-byte lines of each field.  Every thread of the
// block calls this (two barriers).  FUSE: the second sweep (kp.fuseFinal).
template <class R, int L, bool PP, class T, bool FUSE>
__device__ __forceinline__ void sweepYStaged(const KPT<R>& kp, long c, long i, long j, long a, long bx, long by,
                                             int slot, int wave, int lane, bool output, bool hasM, bool hasP,
                                             double vm, double vp, double Lm, double Lp, const T* phiM,
                                             const T* phiP, const double* dfv, bool pb, double* stage,
                                             const double* Uown) {
  constexpr int NC = kNC<L>;
  constexpr int M = kMat<L>;
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const long nx = o.nx, ny = o.ny;
  constexpr int W = kYWaves<R, PP>;
  constexpr int TW = W * kYCols;    // tile width (columns)
  constexpr int TILE = kYOut * TW;  // cells per tile, fewer than the block's threads
  const bool withG = (o.methode == 0 || o.methode == 1);
  const OutStage<TILE> st = OutStage<TILE>::template make<R, L, FUSE>(stage, withG, pb);
  if (output) {
    CellOut<R, L> out;
    double Unew[NC];
    sweepUnew<R, L, PP, false, T>(kp, c, i, j, a, hasM, hasP, vm, vp, Lm, Lp, phiM, phiP, Unew, out.G, Uown);
    if constexpr (FUSE) {
      finalCompute<R, M>(kp, c, Unew, out.fin);
    } else {
#pragma unroll
      for (int q = 0; q < NC; ++q) out.U[q] = Unew[q];
    }
#pragma unroll
    for (int k = 0; k < M; ++k) out.DF[k] = dfv[k];
    st.template put<R, L, FUSE>(out, (slot - kHalo) * TW + wave * kYCols + lane % kYCols, withG, pb);
  }
  __syncthreads();
  const int e = threadIdx.x;
  if (e < TILE) {
    const long row = by * kYOut + e / TW, col = bx * TW + e % TW;
    if (row < ny && col < nx) {
      CellOut<R, L> out;
      st.template get<R, L, FUSE>(out, e, withG, pb);
      const long cc = row * nx + col;
      storeCellOut<R, L, FUSE>(kp, cc, cc, out, withG);  // the minus face of a y-cell has the cell's index
      if (pb) {
#pragma unroll
        for (int k = 0; k < M; ++k) at(f.DF[k], cc) = out.DF[k];
      }
    }
  }
  __syncthreads();  // the batch kernels reuse the tile for their next block
}

// b: linear block index.  Along x the strips of all rows are numbered
// row-major and each wave of the block takes the next one; along y a block
// covers 16 columns x 12 rows, blocks ordered column-group-major per row band.
// L: the remap-vector layout (kMat, kNC; VNR or Eucclhyd).
template <class R, int L, bool PP, bool XD, bool LAG>
__device__ __forceinline__ void sweepWaveBody(const KPT<R>& kp, unsigned b) {
  // type of the slope / limiter / PenteBorne arithmetic
  using T = typename std::conditional<kMixed<R> && !PP, float, double>::type;
  constexpr int NC = kNC<L>;
  constexpr int M = kMat<L>;
  constexpr int S = XD ? 1 : kYCols;  // lane distance of the along-sweep neighbour
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const long nx = o.nx, ny = o.ny;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;

  // cell of this lane: along index a (with na cells), across index t
  long a, t, na, nt;
  long bx = 0, by = 0;  // y: block column group and row band
  int slot;  // position along the sweep inside the wave strip (0 and last: halo)
  if (XD) {
    const long nseg = (nx + kXOut - 1) / kXOut;  // strips per row
    const long g = (long)b * 4 + wave;
    const long seg = g % nseg;
    slot = lane;
    a = seg * kXOut - kHalo + lane;
    t = g / nseg;  // rows beyond ny have no cells
    na = nx, nt = ny;
  } else {
    constexpr int W = kYWaves<R, PP>;
    const long gx = (nx + W * kYCols - 1) / (W * kYCols);
    bx = b % gx, by = b / gx;
    slot = lane / kYCols;
    a = by * kYOut - kHalo + slot;
    t = bx * (W * kYCols) + wave * kYCols + (lane % kYCols);
    na = ny, nt = nx;
  }
  const bool inA = (a >= 0 && a < na), inT = (t >= 0 && t < nt);
  const bool in = inA && inT;
  const long j = XD ? t : a, i = XD ? a : t;
  const long c = j * nx + i;
  const int lastSlot = XD ? kWave - 1 : kYRows - 1;
  const bool output = in && slot >= kHalo && slot <= lastSlot - kHalo;

  double* const* Uin = (kp.sweep == 1) ? f.UL : f.U1;
  // dt1: the time step that moved the nodes to their end-of-Lagrange
  // positions (VNR: dt^{n+1}, updatePosition; Eucclhyd: dt^n,
  // eucclhyd_remap/PhaseLagrange.cc:573-583)
  const double dtn = f.sc->dt[kp.cur], dt1 = f.sc->dt[kEuc<L> ? kp.cur : kp.nxt];
  const bool needH = PP || o.limiter >= MHY_LIM_MINMODG;
  const bool mixteOpt = PP && o.pp_mixte == 1;
  const bool pb = PP && o.methode == 2;

  // LAG: the cell's remap vector at the sweep start, formed by the Lagrange
  // update (no copy in HBM), kept for the cell's own update at the end.  The
  // one-material FP64 sweeps park it in LDS, one slot per thread, so that it
  // holds no registers through the slope computation (each thread reads back
  // only its own slots); the others keep it in registers.
  constexpr bool kPark = LAG && kParkLag<R, L, PP>;
  constexpr int kThreads = XD ? 256 : 64 * kYWaves<R, PP>;
  __shared__ double uPark[kPark ? NC : 1][kPark ? kThreads : 1];
  double uOwn[NC];
  // one cell of this lane's line at along index b: Phi, centroid, width, flags
  auto loadCell = [&](long b, T* ph, double& x, T& h, int& flg) {
    const long jb = XD ? t : b, ib = XD ? b : t;
    const long cb = jb * nx + ib;
    double u[NC];
    double V, qv = 0.0;
    if constexpr (LAG) {
      static_assert(!kEuc<L>, "the fused Lagrange sweep is the VNR cell update");
      // the Lagrange update of the cell, its remap vector in registers; only
      // an updated cell stores what outlives the step
      LagCell<M> lc;
      lagrangeCellCompute<R, M>(kp, cb, jb, ib, output, lc);
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        u[q] = lc.U[q];
        if constexpr (kPark) uPark[q][threadIdx.x] = lc.U[q];
        else uOwn[q] = lc.U[q];
      }
      V = lc.V, qv = lc.q;
      x = XD ? lc.xc.x : lc.xc.y;
      if (output && o.methode == 1) at(f.rL, cb) = lc.rho;  // A2 dual fluxes
      if (output && PP) at(f.qL, cb) = lc.q;                // PenteBorne Phi of the second sweep
      if (output && kp.geomLoad) at(f.VL, cb) = lc.V, at(f.XcL, cb) = lc.xc.x, at(f.YcL, cb) = lc.xc.y;
    } else {
      // (the VNR PenteBorne Phi takes the pseudo from q_L, not from U)
      const int nload = (PP && !kEuc<L>) ? 3 * M : NC;
#pragma unroll
      for (int q = 0; q < NC; ++q) u[q] = (q < nload) ? at(Uin[q], cb) : 0.0;
      // the Lagrange volume and centroid: stored by the cell update, or
      // formed again from the end-of-Lagrange node positions (bitwise the
      // values the cell update formed)
      if (kEuc<L> || kp.geomLoad) {  // (Eucclhyd: always stored)
        V = at(f.VL, cb);
        x = XD ? at(f.XcL, cb) : at(f.YcL, cb);
      } else {
        D2 xcv;
        lagGeometry(kp, dt1, jb, ib, V, xcv);
        x = XD ? xcv.x : xcv.y;
      }
      if (PP && !kEuc<L>) qv = at(f.qL, cb);
    }
    cellPhi<L, PP, T>(u, PP ? 1.0 : V, qv, ph);
    h = needH ? T(widthW<XD>(kp, dt1, jb, ib)) : T(0.0);
    flg = 0;
    if (mixteOpt) flg = (kp.sweep == 1) ? packFlags(at(f.mixte, cb), at(f.pure, cb)) : packedFlags<L, T>(ph);
  };
  T phi[NC], H = T(0.0);
  double xc = 0.0;
  int fl = 0;
#pragma unroll
  for (int q = 0; q < NC; ++q) phi[q] = T(0.0);
  if (in) loadCell(a, phi, xc, H, fl);

  // neighbours along the sweep from the adjacent lanes (the outermost halo
  // lanes clamp to themselves: their slopes are not used); a missing
  // neighbour takes this cell's values (PhaseRemap1.cc:157-158)
  const int lm = (lane - S >= 0) ? lane - S : lane, lp = (lane + S < kWave) ? lane + S : lane;
  const bool inM = (a >= 1), inP = (a + 1 < na);
  T phm[NC], php[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    const T vm = nbrM<XD>(phi[q], lm), vp = nbrP<XD>(phi[q], lp);
    phm[q] = inM ? vm : phi[q];
    php[q] = inP ? vp : phi[q];
  }
  double xcm = nbrM<XD>(xc, lm), xcp = nbrP<XD>(xc, lp);
  xcm = inM ? xcm : xc, xcp = inP ? xcp : xc;
  T Hm = nbrM<XD>(H, lm), Hp = nbrP<XD>(H, lp);
  Hm = inM ? Hm : H, Hp = inP ? Hp : H;
  int fm = nbrM<XD>(fl, lm), fp = nbrP<XD>(fl, lp);
  fm = inM ? fm : fl, fp = inP ? fp : fl;
  // FP64: absolute centroids (the reference's differences); FP32: offsets
  // from this cell's centroid, formed in FP64
  constexpr bool kF32 = sizeof(T) == 4;
  const T xcmT = kF32 ? T(xcm - xc) : T(xcm), xcT = kF32 ? T(0.0) : T(xc);
  const T xcpT = kF32 ? T(xcp - xc) : T(xcp);
  Nbr<T> n{phm, phi, php, xcmT, xcT, xcpT, inM, inP, H, Hm, Hp, false};
  if (mixteOpt) n.vp = ((fl >> 8) == 0) && fl == fp && fl == fm;
  // divided differences (gradPhiFace): each lane forms its minus face's;
  // its plus face is the next lane's minus face (the same operands, so the
  // same value as forming it here)
  T dM[NC], dP[NC];
  if (o.order > 1) {
#pragma unroll
    for (int q = 0; q < NC; ++q) dM[q] = faceDiffM<XD, T>(n, q);
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const T v = nbrP<XD>(dM[q], lp);
      dP[q] = inP ? v : T(0.0);
    }
  } else {
#pragma unroll
    for (int q = 0; q < NC; ++q) dM[q] = dP[q] = T(0.0);
  }
  n.dM = dM;
  n.dP = dP;

  // normal velocities of the minus / plus faces (PrepareRemap.cc:50-62)
  auto fnv = [&](long b) {
    // the face's second node is a uniform distance from the first (next row
    // along x, next column along y): one 32-bit offset serves both loads
    const long n1 = XD ? j * (nx + 1) + b : b * (nx + 1) + i;
    const R* uu = XD ? f.ux[kp.nxt] : f.uy[kp.nxt];
    return 0.0 + ((0.0 + (double)at(uu, n1)) + (double)at(uu + (XD ? nx + 1 : 1), n1)) * 0.5;
  };
  const double vm = in ? fnv(a) : 0.0;
  double vp;
  if constexpr (XD) {
    // along x a lane's plus face is the next lane's minus face (the same
    // nodes, so the same bits), taken by a DPP move where that cell is in the
    // domain; the wave's last lane keeps its own value there, unused: it is
    // a halo lane, and only its own update reads its plus face
    const double vn = nbrP<XD>(vm, lp);
    vp = (in && a + 1 < na) ? vn : (in ? fnv(a + 1) : 0.0);
  } else {
    vp = in ? fnv(a + 1) : 0.0;
  }
  // Euler lengths of the minus / plus faces (Init.cc:528-533).  A face's two
  // nodes share their along-sweep coordinate, whose difference is exactly 0,
  // so the length depends on the across index only: the plus face's is the
  // minus face's, bit for bit.
  auto faceLen = [&](long b) {
    D2 e1, e2;
    if (XD) e1 = eulerNode(o, j, b), e2 = eulerNode(o, j + 1, b);
    else e1 = eulerNode(o, b, i), e2 = eulerNode(o, b, i + 1);
    return d2norm(d2sub(e2, e1));
  };
  const double Lm = faceLen(a), Lp = Lm;
  const bool hasM = inM, hasP = inP;

  T phiM[NC], phiP[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) phiM[q] = phiP[q] = T(0.0);
  double dfv[M];  // PB dual flux of an output cell
#pragma unroll
  for (int k = 0; k < M; ++k) dfv[k] = 0.0;
  if (!PP) {
    // own limited slope, then the upwind value of each face
    // (UtilesRemap-Impl.h:491-501); back/front cells: x: left/right,
    // y: above/below
    T g[NC], gm[NC], gp[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) g[q] = in ? gradQ<XD, T>(o, n, q) : T(0.0);
#pragma unroll
    for (int q = 0; q < NC; ++q) gm[q] = nbrM<XD>(g[q], lm), gp[q] = nbrP<XD>(g[q], lp);
    // face between along-cells lo (lower index) and hi
    auto faceValue = [&](long b, const T* pLo, const T* gLo, double xLo, const T* pHi, const T* gHi, double xHi,
                         double v, T* out) {
      const double dxl = XD ? xHi - xLo : xLo - xHi;  // front - back
      const bool upLo = XD ? (v * dxl > 0.0) : !(v * dxl > 0.0);
      const double xf = XD ? 0.5 * (o.dx * (double)b + o.dx * (double)b) : 0.5 * (o.dy * (double)b + o.dy * (double)b);
      const T s = T(0.0 + (xf - (upLo ? xLo : xHi)));
#pragma unroll
      for (int q = 0; q < NC; ++q) out[q] = upLo ? pLo[q] + gLo[q] * s : pHi[q] + gHi[q] * s;
    };
    if (hasM) faceValue(a, phm, gm, xcm, phi, g, xc, vm, phiM);
    if (hasP) faceValue(a + 1, phi, g, xc, php, gp, xcp, vp, phiP);
  } else {
    // PenteBorne: phi_f = Av(minus cell) - Ar(plus cell) (PhaseRemap1.cc:361-364)
    T Ar[NC], Av[NC], Dual[M];
#pragma unroll
    for (int q = 0; q < NC; ++q) Ar[q] = Av[q] = T(0.0);
#pragma unroll
    for (int k = 0; k < M; ++k) Dual[k] = T(0.0);
    const bool formed = o.order > 1;  // the fluxes are only formed for order > 1
    if (formed && in) ppFluxes<L, XD, T>(o, n, T(vm), T(vp), T(dtn), pb && output, Ar, Av, Dual);
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const T avm = nbrM<XD>(Av[q], lm), arp = nbrP<XD>(Ar[q], lp);
      if (formed) {
        phiM[q] = hasM ? avm - Ar[q] : T(0.0);
        phiP[q] = hasP ? Av[q] - arp : T(0.0);
      }
    }
    if (pb && output) {
      const double w = (Lm + Lp) * 0.5;
#pragma unroll
      for (int k = 0; k < M; ++k) dfv[k] = formed ? (double)Dual[k] * w : 0.0;
    }
  }
  if constexpr (kPark) {
#pragma unroll
    for (int q = 0; q < NC; ++q) uOwn[q] = uPark[q][threadIdx.x];
  }
  if constexpr (XD) {
    if (!output) return;
    if (pb) {
#pragma unroll
      for (int k = 0; k < M; ++k) at(f.DF[k], c) = dfv[k];
    }
    sweepUpdateCell<R, L, PP, XD, T>(kp, c, i, j, a, hasM, hasP, vm, vp, Lm, Lp, phiM, phiP, LAG ? uOwn : nullptr);
  } else {
    // one buffer for both sweep orders: kStageWords doubles per tile cell
    __shared__ double stage[kStageWords<R, L, PP> * kYOut * kYWaves<R, PP> * kYCols];
    const double* Uown = LAG ? uOwn : nullptr;
    bool fused = false;
    if constexpr (!kEuc<L>) {  // the fused final remap is the VNR one
      if (kp.fuseFinal) {
        sweepYStaged<R, L, PP, T, true>(kp, c, i, j, a, bx, by, slot, wave, lane, output, hasM, hasP, vm, vp, Lm, Lp,
                                        phiM, phiP, dfv, pb, stage, Uown);
        fused = true;
      }
    }
    if (!fused)
      sweepYStaged<R, L, PP, T, false>(kp, c, i, j, a, bx, by, slot, wave, lane, output, hasM, hasP, vm, vp, Lm, Lp,
                                       phiM, phiP, dfv, pb, stage, Uown);
  }
}



// waves per SIMD the y kernels are compiled for: two 8-wave blocks per CU in
// the mixed form's limited-slope sweeps (at most 128 VGPRs)
template <class R, bool PP, bool XD>
constexpr int kMinWaves = (!XD && kMixed<R> && !PP) ? 4 : 1;
// the same for the ensemble kernels, whose three-material form would spill
template <class R, int M, bool PP, bool XD>
constexpr int kMinWavesBatch = (M < 3) ? kMinWaves<R, PP, XD> : 1;
// the plain sweeps of one material in FP64 with limited slopes need 98
// VGPRs unbounded; 96 (5 waves per SIMD instead of 4) costs no spill and runs
// the 4096^2 y sweep 22 % faster.  (The fused Lagrange sweeps would spill.)
template <class R, int L, bool PP, bool XD>
constexpr int kMinWavesPlain = (!kMixed<R> && L == 1 && !PP) ? 5 : kMinWaves<R, PP, XD>;
template <class R, int L, bool PP, bool XD>
constexpr int kMinWavesLag = kParkLag<R, L, PP> ? 5 : kMinWaves<R, PP, XD>;

template <class R, int M, bool PP, bool XD>
__global__ __launch_bounds__((XD ? 256 : 64 * kYWaves<R, PP>))
__attribute__((amdgpu_waves_per_eu(kMinWavesPlain<R, M, PP, XD>))) void k_sweep_wave(KPT<R> kp) {
  sweepWaveBody<R, M, PP, XD, false>(kp, XD ? blockIdx.x : xcdBlock());
}

// the first sweep of a step with the cell part of the Lagrange phase
template <class R, int M, bool PP, bool XD>
__global__ __launch_bounds__((XD ? 256 : 64 * kYWaves<R, PP>))
__attribute__((amdgpu_waves_per_eu(kMinWavesLag<R, M, PP, XD>))) void k_sweep_wave_lag(KPT<R> kp) {
  sweepWaveBody<R, M, PP, XD, true>(kp, XD ? blockIdx.x : xcdBlock());
}

// all problems of an ensemble in one launch (BatchLaunchT, common.hpp)
template <class R, int M, bool PP, bool XD, bool ONE>
__global__ __launch_bounds__((XD ? 256 : 64 * kYWaves<R, PP>))
__attribute__((amdgpu_waves_per_eu(kMinWavesBatch<R, M, PP, XD>))) void k_sweep_wave_batch(BatchLaunchT<R> bl) {
  batchRun<ONE>(bl, [&](const KPT<R>& kp, unsigned local, unsigned, unsigned) {
    sweepWaveBody<R, M, PP, XD, false>(kp, local);
  });
}

#define MHY_SWEEP_WAVE_INST(R, M, PP)                                          \
  template __global__ void k_sweep_wave<R, M, PP, true>(KPT<R>);               \
  template __global__ void k_sweep_wave<R, M, PP, false>(KPT<R>);              \
  template __global__ void k_sweep_wave_lag<R, M, PP, true>(KPT<R>);           \
  template __global__ void k_sweep_wave_lag<R, M, PP, false>(KPT<R>);          \
  template __global__ void k_sweep_wave_batch<R, M, PP, true, false>(BatchLaunchT<R>);  \
  template __global__ void k_sweep_wave_batch<R, M, PP, false, false>(BatchLaunchT<R>); \
  template __global__ void k_sweep_wave_batch<R, M, PP, true, true>(BatchLaunchT<R>);   \
  template __global__ void k_sweep_wave_batch<R, M, PP, false, true>(BatchLaunchT<R>);
#define MHY_SWEEP_WAVE_INST_R(R)    \
  MHY_SWEEP_WAVE_INST(R, 1, false) \
  MHY_SWEEP_WAVE_INST(R, 1, true)  \
  MHY_SWEEP_WAVE_INST(R, 2, false) \
  MHY_SWEEP_WAVE_INST(R, 2, true)  \
  MHY_SWEEP_WAVE_INST(R, 3, false) \
  MHY_SWEEP_WAVE_INST(R, 3, true)
MHY_SWEEP_WAVE_INST_R(double)
MHY_SWEEP_WAVE_INST_R(float)
#undef MHY_SWEEP_WAVE_INST_R
#undef MHY_SWEEP_WAVE_INST

// the Eucclhyd layouts (FP64 only): momentum (4 + m), and kinetic energy
// (8 + m) for a conservative projection
#define MHY_SWEEP_WAVE_EUC(L)                                          \
  template __global__ void k_sweep_wave<double, L, false, true>(KP);  \
  template __global__ void k_sweep_wave<double, L, false, false>(KP); \
  template __global__ void k_sweep_wave<double, L, true, true>(KP);   \
  template __global__ void k_sweep_wave<double, L, true, false>(KP);
MHY_SWEEP_WAVE_EUC(5)
MHY_SWEEP_WAVE_EUC(6)
MHY_SWEEP_WAVE_EUC(7)
MHY_SWEEP_WAVE_EUC(9)
MHY_SWEEP_WAVE_EUC(10)
MHY_SWEEP_WAVE_EUC(11)
#undef MHY_SWEEP_WAVE_EUC

// threads per block of the wave sweeps: 4 waves along x; along y 4, or 8
// in the mixed form's limited-slope sweeps
unsigned sweepWaveBlock(bool xd, bool mixed, bool pp) {
  return xd ? 256u : 64u * ((mixed && !pp) ? kYWaves<float, false> : kYWaves<double, false>);
}

// launch grid of the wave sweeps (blocks of sweepWaveBlock threads)
dim3 sweepWaveGrid(bool xd, long nx, long ny, bool mixed, bool pp) {
  if (xd) {
    const long strips = (nx + kXOut - 1) / kXOut * ny;
    return dim3((unsigned)((strips + 3) / 4));
  }
  const int W = (mixed && !pp) ? kYWaves<float, false> : kYWaves<double, false>;
  const long gx = (nx + W * kYCols - 1) / (W * kYCols), gy = (ny + kYOut - 1) / kYOut;
  return dim3((unsigned)(gx * gy));
}

}  // namespace mhy
