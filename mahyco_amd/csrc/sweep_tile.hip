// NOT BUILT.  A 2048x2048 bench run with these kernels ended with a GPU fault
// whose cause was not found: the same code is bit-exact on the GPU on small
// meshes and clean under a host build with bounds checks at 2048x2048.  The
// dispatched sweeps are the per-cell kernels of sweep_cell.hip.
//
// Remap sweeps (Remap::computeGradPhiFace{1,2}, computeGradPhi{1,2},
// computeUpwindFaceQuantitiesForProjection{1,2}, computeUremap{1,2};
// remap/PhaseRemap1.cc, remap/PhaseRemap2.cc) as two CDNA4 kernels:
//
//  * x sweep: LDS tile of XTB rows x XTA cells plus a 2-cell halo along x.
//    Phase 1 stages Phi, centroids, widths and face velocities; phase 2
//    forms each cell's limited slope (or PenteBorne fluxes) once; phase 3
//    builds the two face fluxes of each cell and updates it.
//  * y sweep: column walker.  One thread owns a column segment and slides a
//    window of rows down it in registers, so every row access is a coalesced
//    load of consecutive columns and no LDS is needed.
//
// Both keep the reference's operation order, so FP64 results are bitwise
// those of the CPU path.  The face fluxes of the sweep are also written per
// face for the nodal (dual) momentum remap.
#include "kernels.hpp"
#include "sweep_math.hpp"

namespace mhy {

using namespace sw;

namespace {

constexpr int XTA = 64, XTB = 4;  // x tile: cells along x, rows
constexpr int YTY = 16;           // y walker: rows per thread

// end-of-Lagrange node position (updatePosition, PhaseLagrange.cc:203-208)
__device__ __forceinline__ D2 lagNode(const KP& kp, double dt1, long j, long i) {
  const long p = j * (kp.o.nx + 1) + i;
  const D2 e = eulerNode(kp.o, j, i);
  return {e.x + kp.f.ux[kp.nxt][p] * dt1, e.y + kp.f.uy[kp.nxt][p] * dt1};
}

// widths across the sweep faces (HvLagrange, PhaseRemap1.cc:57-71, :104-119)
template <bool XD>
__device__ __forceinline__ double cellWidth(const KP& kp, double dt1, long j, long i) {
  double h = 0.;
  if (XD) {  // bottom then top face
    h += 0.5 * d2norm(d2sub(lagNode(kp, dt1, j, i + 1), lagNode(kp, dt1, j, i)));
    h += 0.5 * d2norm(d2sub(lagNode(kp, dt1, j + 1, i + 1), lagNode(kp, dt1, j + 1, i)));
  } else {  // right then left face
    h += 0.5 * d2norm(d2sub(lagNode(kp, dt1, j + 1, i + 1), lagNode(kp, dt1, j, i + 1)));
    h += 0.5 * d2norm(d2sub(lagNode(kp, dt1, j + 1, i), lagNode(kp, dt1, j, i)));
  }
  return h;
}

// face normal velocity (PrepareRemap.cc:50-62): mean of the end-of-Lagrange
// velocities of the two face nodes, projected on the face normal
__device__ __forceinline__ double faceVel(const double* u, long n1, long n2) {
  return 0.0 + ((0.0 + u[n1]) + u[n2]) * 0.5;
}

template <int M, bool PP>
__device__ __forceinline__ void loadPhi(const KP& kp, double* const* Uin, long c, double* phi) {
  constexpr int NC = 3 * M + 1;
  double u[NC];
  const int nload = PP ? 3 * M : NC;
#pragma unroll
  for (int q = 0; q < NC; ++q) u[q] = (q < nload) ? Uin[q][c] : 0.0;
  cellPhi<M, PP>(u, PP ? 1.0 : kp.f.VL[c], PP ? kp.f.qL[c] : 0.0, phi);
}

// flux through one cell edge from the opposite face value
// (remap/BoundaryFluxes.cc:14-120); rule 1: minus face, 2: plus face
template <int M, bool PP>
__device__ __forceinline__ void boundaryFlux(int rule, double v, double L, double dtn, const double* phiOpp,
                                             bool oppOk, double* red) {
  constexpr int NC = 3 * M + 1;
  const double sgn = (rule == 1) ? -1.0 : 1.0;
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    const double ph = oppOk ? phiOpp[q] : 0.0;
    const double Fb = PP ? (sgn * L) * ph : (((sgn * v) * L) * dtn) * ph;
    red[q] = red[q] + Fb;
  }
}

__device__ __forceinline__ bool pureNeighbourhood(int fm, int f0, int fp) {
  return ((f0 >> 8) == 0) && f0 == fp && f0 == fm;
}

}  // namespace

// ------------------------------------------------------------------ x sweep
template <int M, bool PP>
__global__ __launch_bounds__(256) void k_sweep_x(KP kp) {
  constexpr int NC = 3 * M + 1;
  constexpr int W = XTA + 4;
  const Opts& o = kp.o;
  const Fields& f = kp.f;
  const long nx = o.nx, ny = o.ny;
  const long a0 = (long)blockIdx.x * XTA, j0 = (long)blockIdx.y * XTB;
  const int tid = threadIdx.x;
  double* const* Uin = (kp.sweep == 1) ? f.UL : f.U1;
  double* const* Uout = (kp.sweep == 1) ? f.U1 : f.U2;
  const double dtn = f.sc->dt[kp.cur], dt1 = f.sc->dt[kp.nxt];
  const bool needH = PP || o.limiter >= MHY_LIM_MINMODG;
  const bool mixteOpt = PP && o.pp_mixte == 1;
  const bool pb = PP && o.methode == 2;

  __shared__ double sPhi[XTB][W][NC];
  __shared__ double sXc[XTB][W];
  __shared__ double sH[XTB][W];
  __shared__ int sFl[XTB][W];
  __shared__ double sFnv[XTB][XTA + 3];              // faces a0-1 .. a0+XTA+1
  __shared__ double sA[XTB][XTA + 2][NC];            // slope (or Ar) of cells a0-1 .. a0+XTA
  __shared__ double sB[PP ? XTB : 1][PP ? XTA + 2 : 1][NC];  // Av

  // phase 1: stage the cells a0-2 .. a0+XTA+1 and the face velocities
  for (int e = tid; e < XTB * W; e += blockDim.x) {
    const int l = e / W, k = e - l * W;
    const long a = a0 - 2 + k, j = j0 + l;
    if (a < 0 || a >= nx || j >= ny) continue;
    const long c = j * nx + a;
    double phi[NC];
    loadPhi<M, PP>(kp, Uin, c, phi);
#pragma unroll
    for (int q = 0; q < NC; ++q) sPhi[l][k][q] = phi[q];
    sXc[l][k] = f.XcL[c];
    sH[l][k] = needH ? cellWidth<true>(kp, dt1, j, a) : 0.0;
    if (mixteOpt) sFl[l][k] = (kp.sweep == 1) ? packFlags(f.mixte[c], f.pure[c]) : packedFlags<M>(phi);
  }
  for (int e = tid; e < XTB * (XTA + 3); e += blockDim.x) {
    const int l = e / (XTA + 3), k = e - l * (XTA + 3);
    const long b = a0 - 1 + k, j = j0 + l;
    if (b < 0 || b > nx || j >= ny) continue;
    sFnv[l][k] = faceVel(f.ux[kp.nxt], j * (nx + 1) + b, (j + 1) * (nx + 1) + b);
  }
  __syncthreads();

  // phase 2: slopes / PenteBorne fluxes of cells a0-1 .. a0+XTA
  if (o.order > 1 || !PP) {
    for (int e = tid; e < XTB * (XTA + 2); e += blockDim.x) {
      const int l = e / (XTA + 2), k = e - l * (XTA + 2);
      const long a = a0 - 1 + k, j = j0 + l;
      if (a < 0 || a >= nx || j >= ny) continue;
      const int kk = k + 1;  // index in the staged arrays
      const bool inM = (a >= 1), inP = (a + 1 < nx);
      const int km = inM ? kk - 1 : kk, kq = inP ? kk + 1 : kk;
      double phm[NC], ph[NC], php[NC];
#pragma unroll
      for (int q = 0; q < NC; ++q) phm[q] = sPhi[l][km][q], ph[q] = sPhi[l][kk][q], php[q] = sPhi[l][kq][q];
      Nbr n{phm, ph, php, sXc[l][km], sXc[l][kk], sXc[l][kq], inM, inP, sH[l][kk], sH[l][km], sH[l][kq], false};
      if (mixteOpt) n.vp = pureNeighbourhood(sFl[l][km], sFl[l][kk], sFl[l][kq]);
      if (!PP) {
#pragma unroll
        for (int q = 0; q < NC; ++q) sA[l][k][q] = gradQ<true>(o, n, q);
      } else {
        double Ar[NC], Av[NC], Dual[M];
        ppFluxes<M, true>(o, n, sFnv[l][k], sFnv[l][k + 1], dtn, pb, Ar, Av, Dual);
#pragma unroll
        for (int q = 0; q < NC; ++q) sA[l][k][q] = Ar[q], sB[l][k][q] = Av[q];
        if (pb && a >= a0 && a < a0 + XTA) {
          // DualphiFlux scaled by the mean length of the two sweep faces
          const double L = d2norm(d2sub(eulerNode(o, j + 1, a), eulerNode(o, j, a)));
          const double L2 = d2norm(d2sub(eulerNode(o, j + 1, a + 1), eulerNode(o, j, a + 1)));
          const double w = (L + L2) * 0.5;
#pragma unroll
          for (int k2 = 0; k2 < M; ++k2) f.DF[k2][j * nx + a] = Dual[k2] * w;
        }
      }
    }
  } else if (pb) {
    for (int e = tid; e < XTB * XTA; e += blockDim.x) {
      const int l = e / XTA, k = e - l * XTA;
      const long a = a0 + k, j = j0 + l;
      if (a < nx && j < ny)
        for (int k2 = 0; k2 < M; ++k2) f.DF[k2][j * nx + a] = 0.0;
    }
  }
  __syncthreads();

  // phase 3: face fluxes and update of cells a0 .. a0+XTA-1
  for (int e = tid; e < XTB * XTA; e += blockDim.x) {
    const int l = e / XTA, k = e - l * XTA;
    const long a = a0 + k, j = j0 + l;
    if (a >= nx || j >= ny) continue;
    const long c = j * nx + a;
    const bool hasM = (a >= 1), hasP = (a + 1 < nx);
    const int kk = k + 2;  // staged index of the cell; slope index = kk - 1
    const double vm = sFnv[l][k + 1], vp = sFnv[l][k + 2];
    const double L = d2norm(d2sub(eulerNode(o, j + 1, a), eulerNode(o, j, a)));
    const double Lp = d2norm(d2sub(eulerNode(o, j + 1, a + 1), eulerNode(o, j, a + 1)));
    double phiM[NC], phiP[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) phiM[q] = phiP[q] = 0.0;
    if (!PP) {
      // upwind value (UtilesRemap-Impl.h:491-501); x: back = left, front = right
      auto faceValue = [&](long b, int kBack, int kFront, double v, double* out) {
        const double dxl = sXc[l][kFront] - sXc[l][kBack];
        const int up = (v * dxl > 0.0) ? kBack : kFront;
        const double xf = 0.5 * (o.dx * (double)b + o.dx * (double)b);
        const double s = 0.0 + (xf - sXc[l][up]);
#pragma unroll
        for (int q = 0; q < NC; ++q) out[q] = sPhi[l][up][q] + sA[l][up - 1][q] * s;
      };
      if (hasM) faceValue(a, kk - 1, kk, vm, phiM);
      if (hasP) faceValue(a + 1, kk, kk + 1, vp, phiP);
    } else if (o.order > 1) {
      // Av(minus cell) - Ar(plus cell) (PhaseRemap1.cc:361-364)
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        phiM[q] = hasM ? sB[l][k][q] - sA[l][k + 1][q] : 0.0;
        phiP[q] = hasP ? sB[l][k + 1][q] - sA[l][k + 2][q] : 0.0;
      }
    }
    const double coefM = PP ? L : vm * L * dtn;
    const double coefP = PP ? Lp : vp * Lp * dtn;
    double Gm[NC], Gp[NC], red[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      Gm[q] = coefM * phiM[q];
      Gp[q] = coefP * phiP[q];
      red[q] = 0.0;
      if (hasM) red[q] = red[q] + (-Gm[q]);
      if (hasP) red[q] = red[q] + Gp[q];
    }
    if (o.flux_any) {
      int rule = 0;
      if (o.flux_left == 1 && a == 0) rule = 1;
      else if (o.flux_right == 1 && a == nx - 1) rule = 2;
      if (rule == 1) boundaryFlux<M, PP>(1, vm, L, dtn, phiP, hasP, red);
      if (rule == 2) boundaryFlux<M, PP>(2, vp, Lp, dtn, phiM, hasM, red);
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) Uout[q][c] = Uin[q][c] - red[q];
    if (o.methode == 0 || o.methode == 1) {
      const int q0 = (o.methode == 0) ? M : 0;
      const long fm = j * (nx + 1) + a;
#pragma unroll
      for (int k2 = 0; k2 < M; ++k2) f.G[k2][fm] = hasM ? Gm[q0 + k2] : 0.0;
      if (!hasP)
#pragma unroll
        for (int k2 = 0; k2 < M; ++k2) f.G[k2][fm + 1] = 0.0;
    }
  }
}

// ------------------------------------------------------------------ y sweep
namespace {

template <int NC>
struct Row {
  double phi[NC];
  double xc, H;
  int fl;
  bool in;
};

template <int NC>
__device__ __forceinline__ void clampRow(Row<NC>& d, const Row<NC>& s) {
#pragma unroll
  for (int q = 0; q < NC; ++q) d.phi[q] = s.phi[q];
  d.xc = s.xc, d.H = s.H, d.fl = s.fl;
}

}  // namespace

// One thread walks rows y0-2 .. y1-1 of its column.  Entering the step for
// row r it holds rows r and r+1, the slope (or PenteBorne Av flux) of row r
// and the face value of face r; it loads row r+2, forms the slope of row
// r+1, the face value of face r+1 and, for r >= y0, updates cell r.  Face t
// lies between rows t-1 (below, the front cell) and t (above, the back cell)
// (CartesianMesh2D.cc:124-128).
template <int M, bool PP>
__global__ __launch_bounds__(256) void k_sweep_y(KP kp) {
  constexpr int NC = 3 * M + 1;
  const Opts& o = kp.o;
  const Fields& f = kp.f;
  const long nx = o.nx, ny = o.ny;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  const long y0 = (long)blockIdx.y * YTY;
  const long y1 = (y0 + YTY < ny) ? y0 + YTY : ny;
  double* const* Uin = (kp.sweep == 1) ? f.UL : f.U1;
  double* const* Uout = (kp.sweep == 1) ? f.U1 : f.U2;
  const double dtn = f.sc->dt[kp.cur], dt1 = f.sc->dt[kp.nxt];
  const bool needH = PP || o.limiter >= MHY_LIM_MINMODG;
  const bool mixteOpt = PP && o.pp_mixte == 1;
  const bool pb = PP && o.methode == 2;
  const double* uy = f.uy[kp.nxt];
  // the horizontal faces of a column all have the same Euler length
  const double L = d2norm(d2sub(eulerNode(o, 0, i + 1), eulerNode(o, 0, i)));
  auto fnvAt = [&](long b) { return faceVel(uy, b * (nx + 1) + i, b * (nx + 1) + i + 1); };

  auto loadRow = [&](long t, Row<NC>& r) {
    r.in = (t >= 0 && t < ny);
    if (!r.in) {
#pragma unroll
      for (int q = 0; q < NC; ++q) r.phi[q] = 0.0;
      r.xc = r.H = 0.0;
      r.fl = 0;
      return;
    }
    const long c = t * nx + i;
    loadPhi<M, PP>(kp, Uin, c, r.phi);
    r.xc = f.YcL[c];
    r.H = needH ? cellWidth<false>(kp, dt1, t, i) : 0.0;
    r.fl = 0;
    if (mixteOpt) r.fl = (kp.sweep == 1) ? packFlags(f.mixte[c], f.pure[c]) : packedFlags<M>(r.phi);
  };

  Row<NC> r0, r1;      // rows r, r+1
  double carry[NC];    // slope (PenteBorne: Av flux) of row r
  double phiFm[NC];    // face value of face r
  double vm = 0.0;     // normal velocity of face r
#pragma unroll
  for (int q = 0; q < NC; ++q) carry[q] = phiFm[q] = 0.0;
  loadRow(y0 - 2, r0);
  loadRow(y0 - 1, r1);
  for (long r = y0 - 2; r < y1; ++r) {
    Row<NC> r2;
    loadRow(r + 2, r2);
    const double vp = (r + 1 >= 0) ? fnvAt(r + 1) : 0.0;  // faces 0 .. ny exist
    // slope / PenteBorne fluxes of row r+1 (rows r, r+1, r+2)
    double nA[NC], nB[NC];  // slope, or Ar and Av
#pragma unroll
    for (int q = 0; q < NC; ++q) nA[q] = nB[q] = 0.0;
    // missing neighbours take the values of row r+1 (PhaseRemap1.cc:157-158)
    if (!r0.in) clampRow(r0, r1);
    if (!r2.in) clampRow(r2, r1);
    if (r1.in) {
      Nbr n{r0.phi, r1.phi, r2.phi, r0.xc, r1.xc, r2.xc, r0.in, r2.in, r1.H, r0.H, r2.H, false};
      if (mixteOpt) n.vp = pureNeighbourhood(r0.fl, r1.fl, r2.fl);
      if (!PP) {
#pragma unroll
        for (int q = 0; q < NC; ++q) nA[q] = gradQ<false>(o, n, q);
      } else if (o.order > 1) {
        double Dual[M];
        ppFluxes<M, false>(o, n, vp, fnvAt(r + 2), dtn, pb, nA, nB, Dual);
        if (pb && r + 1 >= y0 && r + 1 < y1) {
          const double w = (L + L) * 0.5;
#pragma unroll
          for (int k = 0; k < M; ++k) f.DF[k][(r + 1) * nx + i] = Dual[k] * w;
        }
      } else if (pb && r + 1 >= y0 && r + 1 < y1) {
#pragma unroll
        for (int k = 0; k < M; ++k) f.DF[k][(r + 1) * nx + i] = 0.0;
      }
    }
    // face value of face r+1 (rows r and r+1)
    double phiFp[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) phiFp[q] = 0.0;
    if (r0.in && r1.in) {
      if (!PP) {
        // upwind value (UtilesRemap-Impl.h:491-501): back = row r+1
        const double dxl = r0.xc - r1.xc;
        const bool upBack = (vp * dxl > 0.0);
        const double yf = 0.5 * (o.dy * (double)(r + 1) + o.dy * (double)(r + 1));
        const double s = 0.0 + (yf - (upBack ? r1.xc : r0.xc));
#pragma unroll
        for (int q = 0; q < NC; ++q) phiFp[q] = upBack ? r1.phi[q] + nA[q] * s : r0.phi[q] + carry[q] * s;
      } else if (o.order > 1) {
        // Av(row r) - Ar(row r+1) (PhaseRemap2.cc:361-364)
#pragma unroll
        for (int q = 0; q < NC; ++q) phiFp[q] = carry[q] - nA[q];
      }
    }
    if (r >= y0) {
      const long c = r * nx + i;
      const bool hasM = (r >= 1), hasP = (r + 1 < ny);
      const double coefM = PP ? L : vm * L * dtn;
      const double coefP = PP ? L : vp * L * dtn;
      double red[NC], Gm[NC];
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        Gm[q] = coefM * phiFm[q];
        const double Gp = coefP * phiFp[q];
        red[q] = 0.0;
        if (hasM) red[q] = red[q] + (-Gm[q]);
        if (hasP) red[q] = red[q] + Gp;
      }
      if (o.flux_any) {
        if (o.flux_bottom == 1 && c < nx) boundaryFlux<M, PP>(1, vm, L, dtn, phiFp, hasP, red);
        else if (o.flux_top == 1 && c <= nx * (ny - 1) && c < nx * ny)
          boundaryFlux<M, PP>(2, vp, L, dtn, phiFm, hasM, red);
      }
#pragma unroll
      for (int q = 0; q < NC; ++q) Uout[q][c] = Uin[q][c] - red[q];
      if (o.methode == 0 || o.methode == 1) {
        const int q0 = (o.methode == 0) ? M : 0;
#pragma unroll
        for (int k = 0; k < M; ++k) f.G[k][c] = hasM ? Gm[q0 + k] : 0.0;
        if (!hasP)
#pragma unroll
          for (int k = 0; k < M; ++k) f.G[k][c + nx] = 0.0;
      }
    }
    // advance by one row
    r0 = r1;
    r1 = r2;
    vm = vp;
#pragma unroll
    for (int q = 0; q < NC; ++q) carry[q] = PP ? nB[q] : nA[q], phiFm[q] = phiFp[q];
  }
}

#define MHY_SWEEP_INST(M, PP)                              \
  template __global__ void k_sweep_x<M, PP>(KP);          \
  template __global__ void k_sweep_y<M, PP>(KP);
MHY_SWEEP_INST(1, false)
MHY_SWEEP_INST(1, true)
MHY_SWEEP_INST(2, false)
MHY_SWEEP_INST(2, true)
MHY_SWEEP_INST(3, false)
MHY_SWEEP_INST(3, true)
#undef MHY_SWEEP_INST

// launch configuration of the sweeps
dim3 sweepGrid(bool xd, long nx, long ny) {
  if (xd) return dim3((unsigned)((nx + XTA - 1) / XTA), (unsigned)((ny + XTB - 1) / XTB));
  return dim3((unsigned)((nx + 255) / 256), (unsigned)((ny + YTY - 1) / YTY));
}

}  // namespace mhy
