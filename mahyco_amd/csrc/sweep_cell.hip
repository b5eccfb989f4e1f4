// Remap sweeps, one thread per cell (Remap::computeGradPhiFace{1,2},
// computeGradPhi{1,2}, computeUpwindFaceQuantitiesForProjection{1,2},
// computeUremap{1,2}; remap/PhaseRemap1.cc, remap/PhaseRemap2.cc).
//
// Each thread reads the 5 cells a-2 .. a+2 along the sweep (neighbour reads
// hit L1/L2) and forms what its two faces need: the limited slope of the
// upwind cell of each face, or the PenteBorne fluxes Av(a-1), Ar(a), Av(a),
// Ar(a+1).  The material count is a template parameter, so the 5 x (3m+1)
// stencil values live in registers.  Operation order follows the reference,
// so FP64 results are bitwise those of the CPU path.  The face fluxes are
// also written per face for the nodal (dual) momentum remap.
//
// The file also holds the third-order sweep (projectionOrder 3), which reads
// cells a-3 .. a+3 from an LDS tile (k_sweep_o3).
#include "final_math.hpp"
#include "kernels.hpp"
#include "sweep_math.hpp"
#include "sweep_update.hpp"

namespace mhy {

using namespace sw;

namespace {

// width across the sweep faces (HvLagrange, PhaseRemap1.cc:57-71,
// :104-119) from the end-of-Lagrange nodes (updatePosition,
// PhaseLagrange.cc:203-208), in either storage form
template <bool XD, class R>
__device__ __forceinline__ double widthCell(const KPT<R>& kp, double dt1, long j, long i) {
  auto node = [&](long jj, long ii) {
    const long p = jj * (kp.o.nx + 1) + ii;
    const D2 e = eulerNode(kp.o, jj, ii);
    return D2{e.x + (double)kp.f.ux[kp.nxt][p] * dt1, e.y + (double)kp.f.uy[kp.nxt][p] * dt1};
  };
  double h = 0.;
  if (XD) {  // bottom then top face
    h += 0.5 * d2norm(d2sub(node(j, i + 1), node(j, i)));
    h += 0.5 * d2norm(d2sub(node(j + 1, i + 1), node(j + 1, i)));
  } else {  // right then left face
    h += 0.5 * d2norm(d2sub(node(j + 1, i + 1), node(j, i + 1)));
    h += 0.5 * d2norm(d2sub(node(j + 1, i), node(j, i)));
  }
  return h;
}

template <int NC>
struct Slot {
  double phi[NC];
  double xc, H;
  int fl;
};

template <int NC>
__device__ __forceinline__ void copySlot(Slot<NC>& d, const Slot<NC>& s) {
#pragma unroll
  for (int q = 0; q < NC; ++q) d.phi[q] = s.phi[q];
  d.xc = s.xc, d.H = s.H, d.fl = s.fl;
}

// per-lane choice between two slots, by value (keeps everything in registers)
template <int NC>
__device__ __forceinline__ void selectSlot(Slot<NC>& d, bool first, const Slot<NC>& s0, const Slot<NC>& s1) {
#pragma unroll
  for (int q = 0; q < NC; ++q) d.phi[q] = first ? s0.phi[q] : s1.phi[q];
  d.xc = first ? s0.xc : s1.xc;
  d.H = first ? s0.H : s1.H;
  d.fl = first ? s0.fl : s1.fl;
}

__device__ __forceinline__ bool pureNbh(int fm, int f0, int fp) { return ((f0 >> 8) == 0) && f0 == fp && f0 == fm; }

// Third-order face value (remap/UtilesRemap.cc:197-320, ComputeFluxOrdre3).
// Around a face, y[0..5] and h[0..5] are the cell means and widths of the
// three cells on the minus side (farthest first) and the three on the plus
// side, vdt the face normal velocity times dt.  Each interior edge gets the
// value of Colella & Woodward's quartic fit from the two-cell gradients,
// the upwind cell's parabola is averaged over the swept length, and the
// mean is clamped to the intersection of [ym, yp] and the interval the
// upwind slope allows (an empty intersection clamps to 0).  At vdt = 0 the
// extrapolated bound is +-inf or NaN; min and max are the reference's
// (b < a ? b : a, a < b ? b : a), so a NaN bound collapses to the cell mean
// as there.  For vdt < 0 the parabola term enters with the reference's sign
// (Remap::evaluate_fp).
//
// Every quotient of widths alone is formed once per cell and shared by the
// components: in the reference's expressions those quotients are the
// leftmost operands of left-to-right products, so taking them first changes
// no rounding.

// gradient of the cell of width h0 between widths hm and hp
// (Remap::evaluate_grad): g = c0 * (c1 * (yp - y0) + c2 * (y0 - ym))
struct O3Grad {
  double c0, c1, c2;
};
__device__ __forceinline__ O3Grad o3GradCoef(double hm, double h0, double hp) {
  return {h0 / (hm + h0 + hp), (2. * hm + h0) / (h0 + hp), (h0 + 2. * hp) / (hm + h0)};
}
__device__ __forceinline__ double o3Grad(const O3Grad& k, double ym, double y0, double yp) {
  return k.c0 * (k.c1 * (yp - y0) + k.c2 * (y0 - ym));
}

// value at the edge between the cells of widths hm and hp
// (Remap::evaluate_ystar)
struct O3Edge {
  double t1, am, ap, w, s;
};
__device__ __forceinline__ O3Edge o3EdgeCoef(double hmm, double hm, double hp, double hpp) {
  return {(2. * hp * hm) / (hm + hp) * ((hmm + hm) / (2. * hm + hp) - (hpp + hp) / (2. * hp + hm)),
          -hm * (hmm + hm) / (2. * hm + hp), hp * (hp + hpp) / (hm + 2. * hp), hm / (hm + hp),
          1. / (hmm + hm + hp + hpp)};
}
__device__ __forceinline__ double o3Edge(const O3Edge& k, double ym, double yp, double gm, double gp) {
  const double t1 = k.t1 * (yp - ym);
  const double t2 = k.am * gp + k.ap * gm;
  return ym + k.w * (yp - ym) + k.s * (t1 + t2);
}

// the swept-mean factors of a face: x = |vdt|, dx = the upwind width
struct O3Sweep {
  double p1, p2, r;
  bool pos;
};
__device__ __forceinline__ O3Sweep o3SweepCoef(double vdt, double hm, double hp) {
  O3Sweep k;
  k.pos = (vdt >= 0.);
  const double x = k.pos ? vdt : -vdt, dx = k.pos ? hm : hp;
  k.p1 = 0.5 * x / dx;
  k.p2 = 1. - 2. / 3. * x / dx;
  const double nu = k.pos ? vdt / hm : vdt / hp;
  k.r = k.pos ? (1. - nu) / nu : (1. + nu) / nu;
  return k;
}

// y[0..5]: the cell means around the face; eM, e0, eP: the left edge of
// the minus cell, the face, the right edge of the plus cell
__device__ __forceinline__ double o3Flux(const O3Sweep& k, const double* y, double eM, double e0, double eP) {
  const double ym = y[2], yp = y[3];
  double flux;
  if (k.pos) {
    const double six = 6. * (ym - 0.5 * (eM + e0));
    flux = e0 - k.p1 * ((e0 - eM) - k.p2 * six);
  } else {
    const double six = 6. * (yp - 0.5 * (e0 + eP));
    flux = e0 + k.p1 * ((eP - e0) - k.p2 * six);
  }
  const double i1lo = smin(ym, yp), i1hi = smax(ym, yp);
  double i2lo, i2hi;
  if (k.pos) {
    const double ext = ym + k.r * (ym - y[1]);
    i2lo = smin(ym, ext), i2hi = smax(ym, ext);
  } else {
    const double ext = yp - k.r * (yp - y[4]);
    i2lo = smin(yp, ext), i2hi = smax(yp, ext);
  }
  double lo, hi;
  if ((i1hi < i2lo) || (i2hi < i1lo)) lo = 0., hi = 0.;
  else lo = smax(i1lo, i2lo), hi = smin(i1hi, i2hi);
  if (flux < lo) flux = lo;
  if (flux > hi) flux = hi;
  return flux;
}

}  // namespace

template <int M, bool PP, bool XD>
__global__ __launch_bounds__(256) void k_sweep_cell(KP kp) {
  constexpr int NC = 3 * M + 1;
  const Opts& o = kp.o;
  const Fields& f = kp.f;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= kp.ncell) return;
  const long nx = o.nx, ny = o.ny;
  const long j = c / nx, i = c - j * nx;
  const long a = XD ? i : j, na = XD ? nx : ny, stride = XD ? 1 : nx;
  double* const* Uin = (kp.sweep == 1) ? f.UL : f.U1;
  double* const* Uout = (kp.sweep == 1) ? f.U1 : f.U2;
  const double dtn = f.sc->dt[kp.cur], dt1 = f.sc->dt[kp.nxt];
  const bool needH = PP || o.limiter >= MHY_LIM_MINMODG;
  const bool mixteOpt = PP && o.pp_mixte == 1;
  const bool pb = PP && o.methode == 2;

  // stencil slots 0..4 = cells a-2 .. a+2
  Slot<NC> s[5];
  bool in[5];
#pragma unroll
  for (int d = 0; d < 5; ++d) {
    const long b = a + d - 2;
    in[d] = (b >= 0 && b < na);
    if (!in[d]) continue;
    const long cd = c + (d - 2) * stride;
    double u[NC];
    const int nload = PP ? 3 * M : NC;
#pragma unroll
    for (int q = 0; q < NC; ++q) u[q] = (q < nload) ? Uin[q][cd] : 0.0;
    cellPhi<M, PP, double>(u, PP ? 1.0 : f.VL[cd], PP ? f.qL[cd] : 0.0, s[d].phi);
    s[d].xc = XD ? f.XcL[cd] : f.YcL[cd];
    s[d].H = needH ? widthCell<XD>(kp, dt1, XD ? j : b, XD ? b : i) : 0.0;
    s[d].fl = 0;
    if (mixteOpt) s[d].fl = (kp.sweep == 1) ? packFlags(f.mixte[cd], f.pure[cd]) : packedFlags<M, double>(s[d].phi);
  }
  // a missing neighbour takes the values of the cell it neighbours
  // (PhaseRemap1.cc:157-158)
  if (!in[1]) copySlot(s[1], s[2]);
  if (!in[3]) copySlot(s[3], s[2]);
  if (!in[0]) copySlot(s[0], s[1]);
  if (!in[4]) copySlot(s[4], s[3]);

  // neighbourhood of slot d (1..3)
  auto nbr = [&](const Slot<NC>& sm, const Slot<NC>& s0, const Slot<NC>& sp, bool inM, bool inP) {
    Nbr<double> n{sm.phi, s0.phi, sp.phi, sm.xc, s0.xc, sp.xc, inM, inP, s0.H, sm.H, sp.H, false};
    if (mixteOpt) n.vp = pureNbh(sm.fl, s0.fl, sp.fl);
    return n;
  };

  // face normal velocity of face b, between along-cells b-1 and b
  // (PrepareRemap.cc:50-62)
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

  const bool hasM = (a >= 1), hasP = (a + 1 < na);
  const double Lm = faceLen(a), Lp = faceLen(a + 1);
  const double vm = fnv(a), vp = fnv(a + 1);
  double phiM[NC], phiP[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) phiM[q] = phiP[q] = 0.0;
  if (!PP) {
    // upwind value (UtilesRemap-Impl.h:491-501) of a face between slots
    // lo and lo+1; back/front cells: x: left/right, y: above/below
    auto faceValue = [&](long b, const Slot<NC>& sl0, const Slot<NC>& sl1, const Slot<NC>& sl2, const Slot<NC>& sl3,
                         bool in0, bool in1, bool in2, bool in3, double v, double* out) {
      // sl1 = lower along-cell, sl2 = upper along-cell
      const double dxl = XD ? sl2.xc - sl1.xc : sl1.xc - sl2.xc;  // front - back
      const bool upLow = XD ? (v * dxl > 0.0) : !(v * dxl > 0.0);
      Slot<NC> um, u0, up;
      selectSlot(um, upLow, sl0, sl1);
      selectSlot(u0, upLow, sl1, sl2);
      selectSlot(up, upLow, sl2, sl3);
      const bool inM = upLow ? in0 : in1, inP = upLow ? in2 : in3;
      const Nbr<double> n = nbr(um, u0, up, inM, inP);
      const double xf = XD ? 0.5 * (o.dx * (double)b + o.dx * (double)b) : 0.5 * (o.dy * (double)b + o.dy * (double)b);
      const double sx = 0.0 + (xf - u0.xc);
#pragma unroll
      for (int q = 0; q < NC; ++q) out[q] = u0.phi[q] + gradQ<XD, double>(o, n, q) * sx;
    };
    if (hasM) faceValue(a, s[0], s[1], s[2], s[3], in[0], in[1], in[2], in[3], vm, phiM);
    if (hasP) faceValue(a + 1, s[1], s[2], s[3], s[4], in[1], in[2], in[3], in[4], vp, phiP);
  } else if (o.order > 1) {
    // PenteBorne: phi_f = Av(minus cell) - Ar(plus cell) (PhaseRemap1.cc:361-364)
    double Ar0[NC], Av0[NC], Dual[M];
    ppFluxes<M, XD, double>(o, nbr(s[1], s[2], s[3], in[1], in[3]), vm, vp, dtn, pb, Ar0, Av0, Dual);
    if (pb) {
      const double w = (Lm + Lp) * 0.5;
#pragma unroll
      for (int k = 0; k < M; ++k) f.DF[k][c] = Dual[k] * w;
    }
    if (hasM) {
      double ArM[NC], AvM[NC], DM[M];
      ppFluxes<M, XD, double, false, true>(o, nbr(s[0], s[1], s[2], in[0], in[2]), fnv(a - 1), vm, dtn, false, ArM, AvM, DM);
#pragma unroll
      for (int q = 0; q < NC; ++q) phiM[q] = AvM[q] - Ar0[q];
    }
    if (hasP) {
      double ArP[NC], AvP[NC], DP[M];
      ppFluxes<M, XD, double, true, false>(o, nbr(s[2], s[3], s[4], in[2], in[4]), vp, fnv(a + 2), dtn, false, ArP, AvP, DP);
#pragma unroll
      for (int q = 0; q < NC; ++q) phiP[q] = Av0[q] - ArP[q];
    }
  } else if (pb) {
    // the PenteBorne fluxes are only formed for order > 1
#pragma unroll
    for (int k = 0; k < M; ++k) f.DF[k][c] = 0.0;
  }

  // computeUremap (PhaseRemap1.cc:468-503): a cell receives -G through its
  // minus face and +G through its plus face
  const double coefM = PP ? Lm : vm * Lm * dtn;
  const double coefP = PP ? Lp : vp * Lp * dtn;
  double Gm[NC], red[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    Gm[q] = coefM * phiM[q];
    const double Gp = coefP * phiP[q];
    red[q] = 0.0;
    if (hasM) red[q] = red[q] + (-Gm[q]);
    if (hasP) red[q] = red[q] + Gp;
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
        const double ph = oppOk ? (viaMinus ? phiP[q] : phiM[q]) : 0.0;
        const double Fb = PP ? (sgn * Lb) * ph : (((sgn * vb) * Lb) * dtn) * ph;
        red[q] = red[q] + Fb;
      }
    }
  }
  double Unew[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) Unew[q] = Uin[q][c] - red[q];
  if (kp.fuseFinal) {
    finalCell<double, M>(kp, c, Unew);  // remapVariables on the cell, U2 stays in registers
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

// Third-order sweep on an LDS tile: each thread loads one cell (its means
// U/V_L and its width), the block then forms the faces of its inner cells
// from LDS.  Along x a block is 256 cells of a row, 250 updated; along y
// 8 columns x 32 rows, 26 rows updated.  Slots past the domain edge load the
// edge cell, which is what the reference's neighbour lookups return
// (remap/UtilesMesh.cc:5-51).  Arithmetic order follows the reference.
namespace {
constexpr int kO3Block = 256;
constexpr int kO3XOut = kO3Block - 6;
constexpr int kO3YCols = 8;
constexpr int kO3YRows = kO3Block / kO3YCols;
constexpr int kO3YOut = kO3YRows - 6;
}  // namespace

// the faces and the update of the inner cell at along index a, across
// index t, whose tile slot is tid (sY: NC planes of kO3Block means)
template <class R, int L, bool XD>
__device__ __forceinline__ void sweepO3Cell(const KPT<R>& kp, const double* sY, const double* sH, int tid, long a,
                                            long t, long na) {
  constexpr int NC = kNC<L>;
  constexpr int S = XD ? 1 : kO3YCols;  // thread distance of the along-sweep neighbour
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const long nx = o.nx;
  const double dtn = f.sc->dt[kp.cur];
  const long j = XD ? t : a, i = XD ? a : t;
  const long c = j * nx + i;
  // face normal velocity of face b, between along-cells b-1 and b
  // (PrepareRemap.cc:50-62)
  auto fnv = [&](long b) {
    long n1, n2;
    if (XD) n1 = j * (nx + 1) + b, n2 = (j + 1) * (nx + 1) + b;
    else n1 = b * (nx + 1) + i, n2 = b * (nx + 1) + i + 1;
    const R* u = XD ? f.ux[kp.nxt] : f.uy[kp.nxt];
    return 0.0 + ((0.0 + (double)u[n1]) + (double)u[n2]) * 0.5;
  };
  // Euler length of face b (Init.cc:528-533)
  auto faceLen = [&](long b) {
    D2 e1, e2;
    if (XD) e1 = eulerNode(o, j, b), e2 = eulerNode(o, j + 1, b);
    else e1 = eulerNode(o, b, i), e2 = eulerNode(o, b, i + 1);
    return d2norm(d2sub(e2, e1));
  };
  const bool hasM = (a >= 1), hasP = (a + 1 < na);
  const double Lm = faceLen(a), Lp = faceLen(a + 1);
  const double vm = fnv(a), vp = fnv(a + 1);
  double H[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) H[k] = sH[tid + (k - 3) * S];
  O3Grad gk[5];
  O3Edge ek[4];
#pragma unroll
  for (int k = 1; k <= 5; ++k) gk[k - 1] = o3GradCoef(H[k - 1], H[k], H[k + 1]);
#pragma unroll
  for (int k = 1; k <= 4; ++k) ek[k - 1] = o3EdgeCoef(H[k - 1], H[k], H[k + 1], H[k + 2]);
  const O3Sweep swM = o3SweepCoef(vm * dtn, H[2], H[3]), swP = o3SweepCoef(vp * dtn, H[3], H[4]);
  double phiM[NC], phiP[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    double y[7], g[5], e[4];
#pragma unroll
    for (int k = 0; k < 7; ++k) y[k] = sY[q * kO3Block + tid + (k - 3) * S];
#pragma unroll
    for (int k = 1; k <= 5; ++k) g[k - 1] = o3Grad(gk[k - 1], y[k - 1], y[k], y[k + 1]);
#pragma unroll
    for (int k = 1; k <= 4; ++k) e[k - 1] = o3Edge(ek[k - 1], y[k], y[k + 1], g[k - 1], g[k]);
    phiM[q] = hasM ? o3Flux(swM, y, e[0], e[1], e[2]) : 0.0;
    phiP[q] = hasP ? o3Flux(swP, y + 1, e[1], e[2], e[3]) : 0.0;
  }
  sweepUpdateCell<R, L, false, XD, double>(kp, c, i, j, a, hasM, hasP, vm, vp, Lm, Lp, phiM, phiP);
}

// One tile; blk: its linear block index.  R: bulk storage type.  The face
// values are FP64 in both forms (the remap vectors, V_L and the fluxes are
// FP64; the mixed form stores the node velocities the widths come from in
// FP32).  Every thread of the block reaches both barriers (no early
// return), so the ensemble kernels can run several tiles in one workgroup.
// L: the remap-vector layout (kNC; VNR or Eucclhyd).
template <class R, int L, bool XD>
__device__ __forceinline__ void sweepO3Body(const KPT<R>& kp, unsigned blk) {
  constexpr int NC = kNC<L>;
  __shared__ double sY[NC][kO3Block];
  __shared__ double sH[kO3Block];
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const long nx = o.nx, ny = o.ny;
  const int tid = threadIdx.x;
  long a, t, na, nt;
  int slot, last;
  if (XD) {
    const long nseg = (nx + kO3XOut - 1) / kO3XOut;
    const long b = blk;
    t = b / nseg;
    a = (b % nseg) * kO3XOut - 3 + tid;
    slot = tid, last = kO3Block - 1;
    na = nx, nt = ny;
  } else {
    const long gx = (nx + kO3YCols - 1) / kO3YCols;
    const long b = blk;
    slot = tid / kO3YCols, last = kO3YRows - 1;
    a = (b / gx) * kO3YOut - 3 + slot;
    t = (b % gx) * kO3YCols + tid % kO3YCols;
    na = ny, nt = nx;
  }
  const bool inT = (t < nt);
  double* const* Uin = (kp.sweep == 1) ? f.UL : f.U1;
  // the time step of the end-of-Lagrange node positions (VNR: dt^{n+1};
  // Eucclhyd: dt^n, eucclhyd_remap/PhaseLagrange.cc:573-583)
  const double dt1 = f.sc->dt[kEuc<L> ? kp.cur : kp.nxt];
  if (inT) {
    const long b = min(max(a, 0L), na - 1);
    const long j = XD ? t : b, i = XD ? b : t;
    const long cb = j * nx + i;
    const double V = f.VL[cb];
#pragma unroll
    for (int q = 0; q < NC; ++q) sY[q][tid] = Uin[q][cb] / V;
    sH[tid] = widthCell<XD>(kp, dt1, j, i);
  }
  __syncthreads();
  if (inT && a >= 0 && a < na && slot >= 3 && slot <= last - 3) sweepO3Cell<R, L, XD>(kp, sY[0], sH, tid, a, t, na);
  __syncthreads();  // the ensemble kernels reuse the tile for their next block
}

template <class R, int M, bool XD>
__global__ __launch_bounds__(kO3Block) void k_sweep_o3(KPT<R> kp) {
  sweepO3Body<R, M, XD>(kp, blockIdx.x);
}

// all problems of an ensemble in one launch (BatchLaunchT, common.hpp)
template <class R, int M, bool XD, bool ONE>
__global__ __launch_bounds__(kO3Block) void k_sweep_o3_batch(BatchLaunchT<R> bl) {
  batchRun<ONE>(bl, [&](const KPT<R>& kp, unsigned local, unsigned, unsigned) { sweepO3Body<R, M, XD>(kp, local); });
}

dim3 sweepO3Grid(bool xd, long nx, long ny) {
  if (xd) return dim3((unsigned)((nx + kO3XOut - 1) / kO3XOut * ny));
  return dim3((unsigned)((nx + kO3YCols - 1) / kO3YCols * ((ny + kO3YOut - 1) / kO3YOut)));
}

#define MHY_SWEEP_O3_INST(R, M)                                                \
  template __global__ void k_sweep_o3<R, M, true>(KPT<R>);                       \
  template __global__ void k_sweep_o3<R, M, false>(KPT<R>);                      \
  template __global__ void k_sweep_o3_batch<R, M, true, false>(BatchLaunchT<R>);  \
  template __global__ void k_sweep_o3_batch<R, M, false, false>(BatchLaunchT<R>); \
  template __global__ void k_sweep_o3_batch<R, M, true, true>(BatchLaunchT<R>);   \
  template __global__ void k_sweep_o3_batch<R, M, false, true>(BatchLaunchT<R>);
MHY_SWEEP_O3_INST(double, 1)
MHY_SWEEP_O3_INST(double, 2)
MHY_SWEEP_O3_INST(double, 3)
MHY_SWEEP_O3_INST(float, 1)
MHY_SWEEP_O3_INST(float, 2)
MHY_SWEEP_O3_INST(float, 3)
#undef MHY_SWEEP_O3_INST

// the Eucclhyd layouts (FP64 only, no ensembles)
#define MHY_SWEEP_O3_EUC(L)                                  \
  template __global__ void k_sweep_o3<double, L, true>(KP); \
  template __global__ void k_sweep_o3<double, L, false>(KP);
MHY_SWEEP_O3_EUC(5)
MHY_SWEEP_O3_EUC(6)
MHY_SWEEP_O3_EUC(7)
MHY_SWEEP_O3_EUC(9)
MHY_SWEEP_O3_EUC(10)
MHY_SWEEP_O3_EUC(11)
#undef MHY_SWEEP_O3_EUC

#define MHY_SWEEP_CELL_INST(M, PP)                       \
  template __global__ void k_sweep_cell<M, PP, true>(KP); \
  template __global__ void k_sweep_cell<M, PP, false>(KP);
MHY_SWEEP_CELL_INST(1, false)
MHY_SWEEP_CELL_INST(1, true)
MHY_SWEEP_CELL_INST(2, false)
MHY_SWEEP_CELL_INST(2, true)
MHY_SWEEP_CELL_INST(3, false)
MHY_SWEEP_CELL_INST(3, true)
#undef MHY_SWEEP_CELL_INST

}  // namespace mhy
