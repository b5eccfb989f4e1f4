// Nodal (dual) momentum remap of each sweep, the final remap with the EOS,
// and the node-mass / conservation helpers.  The cell sweeps are in
// sweep_wave.hip and sweep_cell.hip.
//
// Layout: structure-of-arrays in HBM (bulk fields of storage type R: FP64 or
// FP32; masses, fluxes and the nodal momentum FP64), one thread per node
// (cell) with the stencil along the sweep read through the L1/L2 caches.
// Arithmetic is FP64, except the velocity / gradient arithmetic of the wave
// kernels in the mixed form (FP32).  The slope limiters are those of the
// cell sweeps (sweep_math.hpp).
#include <type_traits>

#include "common.hpp"
#include "final_math.hpp"
#include "kernels.hpp"
#include "sweep_math.hpp"

namespace mhy {

// Nodal momentum remap of one sweep (remap/DualPhaseRemap1.cc:18-136,
// remap/DualPhaseRemap2.cc:18-127 with remap/UtilesDualRemap.cc and
// UtilesRemap-Impl.h:8-97).  XD: along x (left/right nodes).
template <class R, bool XD>
__device__ __forceinline__ void dualBody(const KPT<R>& kp, unsigned b) {
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const long p = (long)b * blockDim.x + threadIdx.x;
  if (p >= kp.nnode) return;
  const int m = o.nbmat;
  const long nx = o.nx, ny = o.ny;
  const long j = p / (nx + 1), i = p - j * (nx + 1);
  const long a = XD ? i : j, na = XD ? nx : ny;  // node along index in [0, na]
  const long stride = XD ? 1 : nx + 1;
  const double dtn = f.sc->dt[kp.cur];
  const double dt1 = f.sc->dt[kp.nxt];
  const int order2 = o.order - 1;

  // DualPhi and Lagrange coordinate of the nodes a-3..a+3
  double dpx[7], dpy[7], xl[7];
  bool in[7];
  for (int d = 0; d < 7; ++d) {
    const long b = a + d - 3;
    in[d] = (b >= 0 && b <= na);
    if (!in[d]) {
      dpx[d] = dpy[d] = xl[d] = 0.0;
      continue;
    }
    const long q = p + (d - 3) * stride;
    if (kp.sweep == 1) {
      const double ux1 = (double)at(f.ux[kp.nxt], q), uy1 = (double)at(f.uy[kp.nxt], q);
      dpx[d] = ux1, dpy[d] = uy1;
      const long jq = q / (nx + 1), iq = q - jq * (nx + 1);
      const D2 e = eulerNode(o, jq, iq);
      xl[d] = XD ? e.x + ux1 * dt1 : e.y + uy1 * dt1;
    } else {
      // sweep 2 overwrites the velocity in place, so the Lagrange
      // coordinates come from the copy sweep 1 kept
      const double M1 = at(f.UD1[2], q);
      dpx[d] = at(f.UD1[0], q) / M1, dpy[d] = at(f.UD1[1], q) / M1;
      xl[d] = at(f.XL2, q);
    }
  }
  if (kp.sweep == 1) {
    // Lagrange coordinate of this node along the second sweep's axis
    const D2 e = eulerNode(o, j, i);
    at(f.XL2, p) = XD ? e.y + (double)at(f.uy[kp.nxt], p) * dt1 : e.x + (double)at(f.ux[kp.nxt], p) * dt1;
  }
  // divided differences between slots k and k+1 (k = 1..4), each formed
  // once: the plus side of slot d is the minus side of slot d+1
  double D0[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, D1[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (o.order > 1) {
#pragma unroll
    for (int k = 1; k <= 4; ++k) {
      if (!in[k] || !in[k + 1]) continue;
      D0[k] = (dpx[k + 1] - dpx[k]) / (xl[k + 1] - xl[k]);
      D1[k] = (dpy[k + 1] - dpy[k]) / (xl[k + 1] - xl[k]);
    }
  }
  // limited nodal gradient at slot d (2..4) for velocity components 0, 1
  auto gradAt = [&](int d, double* g) {
    g[0] = g[1] = 0.0;
    if (o.order <= 1 || !in[d - 1] || !in[d + 1]) return;
    const double gr0 = D0[d], gl0 = D0[d - 1];
    const double gr1 = D1[d], gl1 = D1[d - 1];
    const double h0 = 0.5 * (xl[d + 1] - xl[d - 1]);
    double hm, hp;
    if (!in[d - 2]) {
      hm = 0., hp = 0.5 * (xl[d + 2] - xl[d]);
    } else if (!in[d + 2]) {
      hp = 0., hm = 0.5 * (xl[d] - xl[d - 2]);
    } else {
      hm = 0.5 * (xl[d] - xl[d - 2]);
      hp = 0.5 * (xl[d + 2] - xl[d]);
    }
    g[0] = sw::limitedGradient<double>(o.limiter, gr0, gl0, dpx[d + 1], dpx[d - 1], h0, hp, hm);
    g[1] = sw::limitedGradient<double>(o.limiter, gr1, gl1, dpy[d + 1], dpy[d - 1], h0, hp, hm);
  };
  double gM[2], g0[2], gP[2];
  gradAt(3, g0);
  gradAt(2, gM);
  gradAt(4, gP);

  // nodal mass fluxes through the dual faces on the minus / plus side
  double plus = 0., minus = 0.;
  {
    // the faces of the sweep orientation touching the node: two faces across
    // the sweep (t-1, t) at along index a, if that index is interior
    const long t = XD ? j : i, nt = XD ? ny : nx;
    const bool interior = (a > 0 && a < na);
    long tf[2];
    int nf = 0;
    if (interior) {
      if (t >= 1) tf[nf++] = t - 1;
      if (t <= nt - 1) tf[nf++] = t;
    }
    // G index of face (along b, across t); cells on both sides of it
    auto gidx = [&](long b, long tt) { return XD ? tt * (nx + 1) + b : b * nx + tt; };
    auto cidx = [&](long b, long tt) { return XD ? tt * nx + b : b * nx + tt; };
    // the along-sweep neighbour's face / cell: a uniform step, folded into the base
    const long dA = XD ? 1 : nx;
    for (int k = 0; k < m; ++k) {
      double pk = 0, mk = 0;
      if (o.methode == 0 || o.methode == 1) {
        for (int s = 0; s < nf; ++s) {
          const long tt = tf[s];
          const double Gf = at(f.G[k], gidx(a, tt));
          double tp = Gf + at(f.G[k] + dA, gidx(a, tt));
          double tm = Gf + at(f.G[k] - dA, gidx(a, tt));
          if (o.methode == 1) tp = at(f.rL, cidx(a, tt)) * tp, tm = at(f.rL - dA, cidx(a, tt)) * tm;
          pk += tp;
          mk += tm;
        }
        const int nbfaces = 2 * nf;
        if (nbfaces != 0) pk /= nbfaces, mk /= nbfaces;
      } else if (o.methode == 2) {
        int nbcell = 0;
        for (int s = 0; s < nf; ++s) {
          const long tt = tf[s];
          pk += at(f.DF[k], cidx(a, tt));
          mk += at(f.DF[k] - dA, cidx(a, tt));
          nbcell++;
        }
        if (nbcell != 0) pk /= nbcell, mk /= nbcell;
      }
      plus += pk;
      minus += mk;
    }
  }
  // upwind velocities (UtilesDualRemap.cc:234-397); a missing neighbour
  // reads as 0 and its mass flux is 0
  auto upwind = [&](bool minusSide, double flux, const double* gNb, double* r) {
    const int dn = minusSide ? 2 : 4;
    const double dNb[2] = {dpx[dn], dpy[dn]}, d0[2] = {dpx[3], dpy[3]};
    const double uN = XD ? dpx[dn] : dpy[dn], u0 = XD ? dpx[3] : dpy[3];
    const double u = minusSide ? 0.5 * (uN + u0) : 0.5 * (u0 + uN);
    const int sign = (u > 0) ? 1 : -1;
    const double h = minusSide ? (xl[3] - xl[dn]) : (xl[dn] - xl[3]);
    const bool fromNb = minusSide ? (flux > 0) : (flux < 0);
    for (int d = 0; d < 2; ++d) {
      const double base = fromNb ? dNb[d] : d0[d];
      const double g = fromNb ? gNb[d] : g0[d];
      r[d] = base + order2 * (0.5 * g * (sign * h - dtn * u));
    }
  };
  double uvm[2], uvp[2];
  upwind(true, minus, gM, uvm);
  upwind(false, plus, gP, uvp);
  double U0, U1, U2;
  if (kp.sweep == 1) {
    const double Mp = at(f.Mp, p);
    U0 = Mp * at(f.ux[kp.nxt], p), U1 = Mp * at(f.uy[kp.nxt], p), U2 = Mp;
  } else {
    U0 = at(f.UD1[0], p), U1 = at(f.UD1[1], p), U2 = at(f.UD1[2], p);
  }
  const double R2 = U2 + minus - plus;
  const double R0 = U0 + minus * uvm[0] - plus * uvp[0];
  const double R1 = U1 + minus * uvm[1] - plus * uvp[1];
  if (kp.sweep == 1) {
    at(f.UD1[0], p) = R0, at(f.UD1[1], p) = R1, at(f.UD1[2], p) = R2;
  } else {
    // remapVariables, nodes (PhaseRemapFinal.cc:207-234)
    at(f.ux[kp.nxt], p) = R0 / R2;
    at(f.uy[kp.nxt], p) = R1 / R2;
  }
}

template <class R, bool XD>
__global__ __launch_bounds__(256) void k_dual(KPT<R> kp) {
  dualBody<R, XD>(kp, blockIdx.x);
}
template <class R, bool XD, bool ONE>
__global__ __launch_bounds__(256) void k_dual_batch(BatchLaunchT<R> bl) {
  batchRun<ONE>(bl, [&](const KPT<R>& kp, unsigned local, unsigned, unsigned) { dualBody<R, XD>(kp, local); });
}

// The same nodal remap with one lane per node and the neighbour values taken
// by wave shuffles.  A wave covers a strip of nodes along the sweep: along x,
// 64 nodes of a node row; along y, 4 node columns x 16 node rows.  A lane
// needs its neighbours' gradients, which need nodes two further away, so the
// 3 strip nodes at each end are halo (58 of 64, resp. 10 of 16 nodes are
// updated).  Each lane computes its own DualPhi, Lagrange coordinate and
// limited gradient once (the per-node kernel above forms three gradients per
// node).
namespace {
constexpr int kDW = 64;
constexpr int kDYCols = 4;
// waves per block along y: 16 (FP64) or 32 (FP32) node columns
template <class R>
constexpr int kDYWaves = kMixed<R> ? 8 : 4;
constexpr int kDYRows = kDW / kDYCols;
constexpr int kDXOut = kDW - 6;
constexpr int kDYOut = kDYRows - 6;
__device__ __forceinline__ double dshfl(double v, int src) { return __shfl(v, src, kDW); }
__device__ __forceinline__ float dshfl(float v, int src) { return __shfl(v, src, kDW); }
// value of the node OFF (|OFF| <= 2) strip positions along the sweep.  Along
// x, DPP moves by one lane (twice for two); along y, a shuffle of OFF *
// kDYCols lanes.  A lane whose source lies outside the wave keeps its own
// value, except that along x a lane one position from the wave's end gets the
// end lane's value for OFF = -2 / +2: both are halo lanes of the strip, whose
// gradients no updated node reads.
template <bool XD, int OFF, class V>
__device__ __forceinline__ V dnbr(V v, int lane) {
  if constexpr (XD) {
    if constexpr (OFF == -2) return dppMove<kDppWaveShr1>(dppMove<kDppWaveShr1>(v));
    else if constexpr (OFF == -1) return dppMove<kDppWaveShr1>(v);
    else if constexpr (OFF == 1) return dppMove<kDppWaveShl1>(v);
    else return dppMove<kDppWaveShl1>(dppMove<kDppWaveShl1>(v));
  } else {
    const int src = lane + OFF * kDYCols;
    return dshfl(v, (src >= 0 && src < kDW) ? src : lane);
  }
}
}  // namespace

// The nodal update of node p (along index a, across index t) from its
// DualPhi and Lagrange coordinate and those of its two along-sweep
// neighbours (dpx, dpy, xl: nodes a-1, a, a+1), its limited gradient g0 and
// the neighbours' gM, gP (zero outside the domain): mass fluxes through the
// dual faces, upwind velocities, and the nodal vector of sweep 1 or the
// velocity of sweep 2.
template <class R, bool XD, class T>
__device__ __forceinline__ void dualNodeFinish(const KPT<R>& kp, long a, long t, long na, long nt, long p,
                                               const T* dpx, const T* dpy, const double* xl, const T* g0,
                                               const T* gM, const T* gP) {
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const int m = o.nbmat;
  const long nx = o.nx;
  const double dtn = f.sc->dt[kp.cur];
  const int order2 = o.order - 1;

  // nodal mass fluxes through the dual faces on the minus / plus side (FP64)
  double plus = 0., minus = 0.;
  {
    const bool interior = (a > 0 && a < na);
    long tf[2];
    int nf = 0;
    if (interior) {
      if (t >= 1) tf[nf++] = t - 1;
      if (t <= nt - 1) tf[nf++] = t;
    }
    auto gidx = [&](long b, long tt) { return XD ? tt * (nx + 1) + b : b * nx + tt; };
    auto cidx = [&](long b, long tt) { return XD ? tt * nx + b : b * nx + tt; };
    // the along-sweep neighbour's face / cell: a uniform step, folded into the base
    const long dA = XD ? 1 : nx;
    for (int k = 0; k < m; ++k) {
      double pk = 0, mk = 0;
      if (o.methode == 0 || o.methode == 1) {
        for (int s2 = 0; s2 < nf; ++s2) {
          const long tt = tf[s2];
          const double Gf = at(f.G[k], gidx(a, tt));
          double tp = Gf + at(f.G[k] + dA, gidx(a, tt));
          double tm = Gf + at(f.G[k] - dA, gidx(a, tt));
          if (o.methode == 1) tp = at(f.rL, cidx(a, tt)) * tp, tm = at(f.rL - dA, cidx(a, tt)) * tm;
          pk += tp;
          mk += tm;
        }
        const int nbfaces = 2 * nf;
        if (nbfaces != 0) pk /= nbfaces, mk /= nbfaces;
      } else if (o.methode == 2) {
        int nbcell = 0;
        for (int s2 = 0; s2 < nf; ++s2) {
          const long tt = tf[s2];
          pk += at(f.DF[k], cidx(a, tt));
          mk += at(f.DF[k] - dA, cidx(a, tt));
          nbcell++;
        }
        if (nbcell != 0) pk /= nbcell, mk /= nbcell;
      }
      plus += pk;
      minus += mk;
    }
  }
  // upwind velocities (UtilesDualRemap.cc:234-397); a missing neighbour
  // reads as 0 and its mass flux is 0
  const T dtnT = T(dtn);
  auto upwind = [&](bool minusSide, double flux, const T* gNb, double* r) {
    const int dn = minusSide ? 0 : 2;
    const T dNb[2] = {dpx[dn], dpy[dn]}, d0[2] = {dpx[1], dpy[1]};
    const T uN = XD ? dpx[dn] : dpy[dn], u0 = XD ? dpx[1] : dpy[1];
    const T u = minusSide ? T(0.5) * (uN + u0) : T(0.5) * (u0 + uN);
    const int sign = (u > T(0)) ? 1 : -1;
    const T h = T(minusSide ? (xl[1] - xl[dn]) : (xl[dn] - xl[1]));
    const bool fromNb = minusSide ? (flux > 0) : (flux < 0);
    for (int d = 0; d < 2; ++d) {
      const T base = fromNb ? dNb[d] : d0[d];
      const T g = fromNb ? gNb[d] : g0[d];
      r[d] = (double)(base + T(order2) * (T(0.5) * g * (T(sign) * h - dtnT * u)));
    }
  };
  double uvm[2], uvp[2];
  upwind(true, minus, gM, uvm);
  upwind(false, plus, gP, uvp);
  double U0, U1, U2;
  if (kp.sweep == 1) {
    const double Mp = at(f.Mp, p);
    U0 = Mp * at(f.ux[kp.nxt], p), U1 = Mp * at(f.uy[kp.nxt], p), U2 = Mp;
  } else {
    U0 = at(f.UD1[0], p), U1 = at(f.UD1[1], p), U2 = at(f.UD1[2], p);
  }
  const double R2 = U2 + minus - plus;
  const double R0 = U0 + minus * uvm[0] - plus * uvp[0];
  const double R1 = U1 + minus * uvm[1] - plus * uvp[1];
  if (kp.sweep == 1) {
    at(f.UD1[0], p) = R0, at(f.UD1[1], p) = R1, at(f.UD1[2], p) = R2;
  } else {
    // remapVariables, nodes (PhaseRemapFinal.cc:207-234)
    at(f.ux[kp.nxt], p) = R0 / R2;
    at(f.uy[kp.nxt], p) = R1 / R2;
  }
}

// b: linear block index (strips numbered row-major along x, as the sweeps)
template <class R, bool XD>
__device__ __forceinline__ void dualWaveBody(const KPT<R>& kp, unsigned b) {
  // velocity / gradient arithmetic: FP32 in the mixed form; masses, momenta
  // and coordinates stay FP64
  using T = typename std::conditional<kMixed<R>, float, double>::type;
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const long nx = o.nx, ny = o.ny;
  const int lane = threadIdx.x & (kDW - 1), wave = threadIdx.x >> 6;
  long a, t, na, nt;
  int slot, last;
  if (XD) {
    const long nseg = (nx + 1 + kDXOut - 1) / kDXOut;  // strips per node row
    const long g = (long)b * 4 + wave;
    slot = lane, last = kDW - 1;
    a = (g % nseg) * kDXOut - 3 + lane;
    t = g / nseg;
    na = nx, nt = ny;  // node along index in [0, na], across in [0, nt]
  } else {
    constexpr int W = kDYWaves<R>;
    const long gx = (nx + 1 + W * kDYCols - 1) / (W * kDYCols);
    const long bx = b % gx, by = b / gx;
    slot = lane / kDYCols, last = kDYRows - 1;
    a = by * kDYOut - 3 + slot;
    t = bx * (W * kDYCols) + wave * kDYCols + (lane % kDYCols);
    na = ny, nt = nx;
  }
  const bool in = (a >= 0 && a <= na && t >= 0 && t <= nt);
  const long j = XD ? t : a, i = XD ? a : t;
  const long p = j * (nx + 1) + i;
  const bool output = in && slot >= 3 && slot <= last - 3;
  const double dt1 = f.sc->dt[kp.nxt];

  // own DualPhi and Lagrange coordinate (coordinates stay FP64: only their
  // differences enter T arithmetic)
  T dpx0 = T(0.0), dpy0 = T(0.0);
  double xl0 = 0.0;
  if (in) {
    if (kp.sweep == 1) {
      const double ux1 = (double)at(f.ux[kp.nxt], p), uy1 = (double)at(f.uy[kp.nxt], p);
      dpx0 = T(ux1), dpy0 = T(uy1);
      const D2 e = eulerNode(o, j, i);
      xl0 = XD ? e.x + ux1 * dt1 : e.y + uy1 * dt1;
      // Lagrange coordinate along the second sweep's axis (the second sweep
      // overwrites the velocity in place)
      if (output) at(f.XL2, p) = XD ? e.y + uy1 * dt1 : e.x + ux1 * dt1;
    } else {
      const double M1 = at(f.UD1[2], p);
      dpx0 = T(at(f.UD1[0], p) / M1), dpy0 = T(at(f.UD1[1], p) / M1);
      xl0 = at(f.XL2, p);
    }
  }
  // nodes a-2 .. a+2 (slots 1..5 of the reference stencil a-3 .. a+3)
  T dpx[7], dpy[7];
  double xl[7];
  bool inS[7];
  dpx[1] = dnbr<XD, -2>(dpx0, lane), dpy[1] = dnbr<XD, -2>(dpy0, lane), xl[1] = dnbr<XD, -2>(xl0, lane);
  dpx[2] = dnbr<XD, -1>(dpx0, lane), dpy[2] = dnbr<XD, -1>(dpy0, lane), xl[2] = dnbr<XD, -1>(xl0, lane);
  dpx[4] = dnbr<XD, 1>(dpx0, lane), dpy[4] = dnbr<XD, 1>(dpy0, lane), xl[4] = dnbr<XD, 1>(xl0, lane);
  dpx[5] = dnbr<XD, 2>(dpx0, lane), dpy[5] = dnbr<XD, 2>(dpy0, lane), xl[5] = dnbr<XD, 2>(xl0, lane);
#pragma unroll
  for (int d = 1; d <= 5; ++d) {
    if (d == 3) continue;
    const long b = a + (d - 3);
    inS[d] = (b >= 0 && b <= na);
    if (!inS[d]) dpx[d] = dpy[d] = T(0.0), xl[d] = 0.0;
  }
  dpx[3] = dpx0, dpy[3] = dpy0, xl[3] = xl0;
  inS[0] = (a - 3 >= 0), inS[6] = (a + 3 <= na);
  dpx[0] = dpy[0] = dpx[6] = dpy[6] = T(0.0);
  xl[0] = xl[6] = 0.0;

  // own limited nodal gradient (UtilesRemap-Impl.h:8-97).  Each lane forms
  // the divided differences towards its minus neighbour; those towards the
  // plus neighbour are the next lane's (the same operands).
  T gl0 = T(0.0), gl1 = T(0.0);
  if (o.order > 1) {
    gl0 = (dpx[3] - dpx[2]) / T(xl[3] - xl[2]);
    gl1 = (dpy[3] - dpy[2]) / T(xl[3] - xl[2]);
  }
  const T gr0 = dnbr<XD, 1>(gl0, lane), gr1 = dnbr<XD, 1>(gl1, lane);
  T g0[2] = {T(0.0), T(0.0)};
  if (in && o.order > 1 && inS[2] && inS[4]) {
    const int d = 3;
    const T h0 = T(0.5 * (xl[d + 1] - xl[d - 1]));
    T hm, hp;
    if (!inS[d - 2]) {
      hm = T(0.), hp = T(0.5 * (xl[d + 2] - xl[d]));
    } else if (!inS[d + 2]) {
      hp = T(0.), hm = T(0.5 * (xl[d] - xl[d - 2]));
    } else {
      hm = T(0.5 * (xl[d] - xl[d - 2]));
      hp = T(0.5 * (xl[d + 2] - xl[d]));
    }
    g0[0] = sw::limitedGradient<T>(o.limiter, gr0, gl0, dpx[d + 1], dpx[d - 1], h0, hp, hm);
    g0[1] = sw::limitedGradient<T>(o.limiter, gr1, gl1, dpy[d + 1], dpy[d - 1], h0, hp, hm);
  }
  T gM[2], gP[2];
  gM[0] = dnbr<XD, -1>(g0[0], lane), gM[1] = dnbr<XD, -1>(g0[1], lane);
  gP[0] = dnbr<XD, 1>(g0[0], lane), gP[1] = dnbr<XD, 1>(g0[1], lane);
  if (!output) return;
  // a neighbour outside the domain has a zero gradient
  if (!inS[2]) gM[0] = gM[1] = T(0.0);
  if (!inS[4]) gP[0] = gP[1] = T(0.0);
  dualNodeFinish<R, XD, T>(kp, a, t, na, nt, p, &dpx[2], &dpy[2], &xl[2], g0, gM, gP);
}

// Nodal remap along y with one lane per node column: a wave covers 64
// consecutive node columns and walks down a band of kDWalkRows node rows,
// plus 3 halo rows above and below, so every row access is a whole 64-node
// row segment.  Each iteration loads one new row r into a 7-row register
// window, forms the divided difference between rows r-1 and r, the limited
// gradient of row r-2 and the update of row r-3.  Every value is formed
// from the same operands as in the wave-strip kernel, so results are
// bitwise the same.
namespace {
constexpr int kDWalkRows = 32;
}  // namespace

template <class R>
__device__ __forceinline__ void dualWalkBody(const KPT<R>& kp, long item) {
  using T = typename std::conditional<kMixed<R>, float, double>::type;
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const long nx = o.nx, na = o.ny, nt = o.nx;
  const long gx = (nx + 1 + kDW - 1) / kDW;
  const long t = (item % gx) * kDW + (threadIdx.x & (kDW - 1));
  const long a0 = (item / gx) * kDWalkRows;                               // first row updated
  const long a1 = (a0 + kDWalkRows - 1 < na) ? a0 + kDWalkRows - 1 : na;  // last row updated
  const bool inT = (t <= nt);
  const double dt1 = f.sc->dt[kp.nxt];
  const bool ord = o.order > 1;
  // window slot k holds row r-6+k; gradients of rows r-4, r-3, r-2; divided
  // differences of rows r-2, r-1, r (towards the row above)
  T wx[7], wy[7], g0x[3], g0y[3], glx[3], gly[3];
  double wl[7];
  bool wa[7];  // row inside [0, na]
#pragma unroll
  for (int k = 0; k < 7; ++k) wx[k] = wy[k] = T(0.0), wl[k] = 0.0, wa[k] = false;
#pragma unroll
  for (int k = 0; k < 3; ++k) g0x[k] = g0y[k] = glx[k] = gly[k] = T(0.0);
  for (long r = a0 - 3; r <= a1 + 3; ++r) {
#pragma unroll
    for (int k = 0; k < 6; ++k) wx[k] = wx[k + 1], wy[k] = wy[k + 1], wl[k] = wl[k + 1], wa[k] = wa[k + 1];
#pragma unroll
    for (int k = 0; k < 2; ++k) g0x[k] = g0x[k + 1], g0y[k] = g0y[k + 1], glx[k] = glx[k + 1], gly[k] = gly[k + 1];
    // own DualPhi and Lagrange coordinate of row r
    const bool rin = (r >= 0 && r <= na);
    T dx = T(0.0), dy = T(0.0);
    double xl = 0.0;
    if (rin && inT) {
      const long p = r * (nx + 1) + t;
      if (kp.sweep == 1) {
        const double ux1 = (double)at(f.ux[kp.nxt], p), uy1 = (double)at(f.uy[kp.nxt], p);
        dx = T(ux1), dy = T(uy1);
        const D2 e = eulerNode(o, r, t);
        xl = e.y + uy1 * dt1;
        if (r >= a0 && r <= a1) at(f.XL2, p) = e.x + ux1 * dt1;  // the second sweep's axis
      } else {
        const double M1 = at(f.UD1[2], p);
        dx = T(at(f.UD1[0], p) / M1), dy = T(at(f.UD1[1], p) / M1);
        xl = at(f.XL2, p);
      }
    }
    wx[6] = dx, wy[6] = dy, wl[6] = xl, wa[6] = rin;
    // divided difference of row r towards row r-1 (zeros outside the domain)
    glx[2] = gly[2] = T(0.0);
    if (ord) {
      glx[2] = (wx[6] - wx[5]) / T(wl[6] - wl[5]);
      gly[2] = (wy[6] - wy[5]) / T(wl[6] - wl[5]);
    }
    // limited gradient of row c = r-2 (window slot 4)
    g0x[2] = g0y[2] = T(0.0);
    if (wa[4] && inT && ord && wa[3] && wa[5]) {
      const T h0 = T(0.5 * (wl[5] - wl[3]));
      T hm, hp;
      if (!wa[2]) {
        hm = T(0.), hp = T(0.5 * (wl[6] - wl[4]));
      } else if (!wa[6]) {
        hp = T(0.), hm = T(0.5 * (wl[4] - wl[2]));
      } else {
        hm = T(0.5 * (wl[4] - wl[2]));
        hp = T(0.5 * (wl[6] - wl[4]));
      }
      g0x[2] = sw::limitedGradient<T>(o.limiter, glx[1], glx[0], wx[5], wx[3], h0, hp, hm);
      g0y[2] = sw::limitedGradient<T>(o.limiter, gly[1], gly[0], wy[5], wy[3], h0, hp, hm);
    }
    // update of row a = r-3 (window slot 3)
    const long a = r - 3;
    if (a >= a0 && a <= a1 && wa[3] && inT) {
      const T dpx[3] = {wx[2], wx[3], wx[4]}, dpy[3] = {wy[2], wy[3], wy[4]};
      const double xls[3] = {wl[2], wl[3], wl[4]};
      const T g0[2] = {g0x[1], g0y[1]};
      T gM[2] = {g0x[0], g0y[0]}, gP[2] = {g0x[2], g0y[2]};
      if (!wa[2]) gM[0] = gM[1] = T(0.0);
      if (!wa[4]) gP[0] = gP[1] = T(0.0);
      dualNodeFinish<R, false, T>(kp, a, t, na, nt, a * (nx + 1) + t, dpx, dpy, xls, g0, gM, gP);
    }
  }
}

template <class R>
__global__ __launch_bounds__(256) void k_dual_walk(KPT<R> kp) {
  const long gx = (kp.o.nx + 1 + kDW - 1) / kDW;
  const long bands = (kp.o.ny + 1 + kDWalkRows - 1) / kDWalkRows;
  const long item = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (item >= gx * bands) return;  // the whole wave
  dualWalkBody<R>(kp, item);
}

dim3 dualWalkGrid(long nx, long ny) {
  const long gx = (nx + 1 + kDW - 1) / kDW, bands = (ny + 1 + kDWalkRows - 1) / kDWalkRows;
  return dim3((unsigned)((gx * bands + 3) / 4));
}

// Nodal remap along y on an LDS tile: a 256-thread block covers 16 node
// columns and ROWS node rows (3 halo rows above and below, ROWS - 6 rows
// updated), each thread ROWS / 16 rows of one column, so that every load and
// store instruction of a wave touches 4 whole row segments of 16 nodes (128
// bytes in FP64).  The tuner picks the tile height: 48 rows keep the tile at
// 30 KB of LDS and 86 VGPRs, 5 blocks per CU (on a 4096^2 mesh 64 rows, 40
// KB and 4 blocks, are 20 % slower, 32 rows, with more halo, 5 % slower).
// The block loads DualPhi and the Lagrange coordinate of the tile into LDS,
// forms each row's limited gradient from its neighbours there, and updates
// the inner rows; every value comes from the same operands as in the
// wave-strip kernel, so results are bitwise the same.
namespace {
constexpr int kDTThreads = 256;                 // threads per tile (512 and 1024: no faster)
constexpr int kDTCols = 16;                     // node columns per tile
}  // namespace

template <class R, int ROWS>
__device__ __forceinline__ void dualTileBody(const KPT<R>& kp, unsigned b) {
  static_assert(ROWS % (kDTThreads / kDTCols) == 0 && ROWS > 6, "whole passes of 16 rows");
  constexpr int kDTRows = ROWS;
  constexpr int kDTRowsPer = kDTRows * kDTCols / kDTThreads;  // rows per thread
  constexpr int kDTOut = kDTRows - 6;                         // rows updated
  using T = typename std::conditional<kMixed<R>, float, double>::type;
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const long nx = o.nx, na = o.ny, nt = o.nx;
  const long gx = (nx + 1 + kDTCols - 1) / kDTCols;
  const long bx = b % gx, by = b / gx;
  const int col = threadIdx.x % kDTCols, rg = threadIdx.x / kDTCols;
  const long t = bx * kDTCols + col;
  const long r0 = by * kDTOut - 3;  // node row of tile row 0
  const bool inT = (t <= nt);
  const double dt1 = f.sc->dt[kp.nxt];
  __shared__ T sdx[kDTRows][kDTCols], sdy[kDTRows][kDTCols], sgx[kDTRows][kDTCols], sgy[kDTRows][kDTCols];
  __shared__ double sxl[kDTRows][kDTCols];
  // tile row of the k-th row of this thread: 16 consecutive rows per pass
  auto tileRow = [&](int k) { return rg + k * (kDTThreads / kDTCols); };
  auto inRow = [&](int L) { return (r0 + L >= 0) && (r0 + L <= na); };
  // own DualPhi and Lagrange coordinate of the tile's nodes
#pragma unroll
  for (int k = 0; k < kDTRowsPer; ++k) {
    const int L = tileRow(k);
    const long a = r0 + L;
    T dx = T(0.0), dy = T(0.0);
    double xl = 0.0;
    if (inRow(L) && inT) {
      const long p = a * (nx + 1) + t;
      if (kp.sweep == 1) {
        const double ux1 = (double)at(f.ux[kp.nxt], p), uy1 = (double)at(f.uy[kp.nxt], p);
        dx = T(ux1), dy = T(uy1);
        const D2 e = eulerNode(o, a, t);
        xl = e.y + uy1 * dt1;
        // Lagrange coordinate along the second sweep's axis (the second sweep
        // overwrites the velocity in place)
        if (L >= 3 && L < 3 + kDTOut) at(f.XL2, p) = e.x + ux1 * dt1;
      } else {
        const double M1 = at(f.UD1[2], p);
        dx = T(at(f.UD1[0], p) / M1), dy = T(at(f.UD1[1], p) / M1);
        xl = at(f.XL2, p);
      }
    }
    sdx[L][col] = dx, sdy[L][col] = dy, sxl[L][col] = xl;
  }
  __syncthreads();
  // limited gradients of the rows the updates need (tile rows 2 .. ROWS - 3)
#pragma unroll
  for (int k = 0; k < kDTRowsPer; ++k) {
    const int L = tileRow(k);
    if (L < 2 || L >= kDTRows - 2) continue;
    T g0x = T(0.0), g0y = T(0.0);
    if (inT && o.order > 1 && inRow(L) && inRow(L - 1) && inRow(L + 1)) {
      // divided differences towards the row above, at rows L and L+1
      const T glx = (sdx[L][col] - sdx[L - 1][col]) / T(sxl[L][col] - sxl[L - 1][col]);
      const T gly = (sdy[L][col] - sdy[L - 1][col]) / T(sxl[L][col] - sxl[L - 1][col]);
      const T grx = (sdx[L + 1][col] - sdx[L][col]) / T(sxl[L + 1][col] - sxl[L][col]);
      const T gry = (sdy[L + 1][col] - sdy[L][col]) / T(sxl[L + 1][col] - sxl[L][col]);
      const T h0 = T(0.5 * (sxl[L + 1][col] - sxl[L - 1][col]));
      T hm, hp;
      if (!inRow(L - 2)) {
        hm = T(0.), hp = T(0.5 * (sxl[L + 2][col] - sxl[L][col]));
      } else if (!inRow(L + 2)) {
        hp = T(0.), hm = T(0.5 * (sxl[L][col] - sxl[L - 2][col]));
      } else {
        hm = T(0.5 * (sxl[L][col] - sxl[L - 2][col]));
        hp = T(0.5 * (sxl[L + 2][col] - sxl[L][col]));
      }
      g0x = sw::limitedGradient<T>(o.limiter, grx, glx, sdx[L + 1][col], sdx[L - 1][col], h0, hp, hm);
      g0y = sw::limitedGradient<T>(o.limiter, gry, gly, sdy[L + 1][col], sdy[L - 1][col], h0, hp, hm);
    }
    sgx[L][col] = g0x, sgy[L][col] = g0y;
  }
  __syncthreads();
  // updates of the inner rows
#pragma unroll
  for (int k = 0; k < kDTRowsPer; ++k) {
    const int L = tileRow(k);
    if (L < 3 || L >= 3 + kDTOut) continue;
    const long a = r0 + L;
    if (!inT || !inRow(L)) continue;
    const T dpx[3] = {sdx[L - 1][col], sdx[L][col], sdx[L + 1][col]};
    const T dpy[3] = {sdy[L - 1][col], sdy[L][col], sdy[L + 1][col]};
    const double xls[3] = {sxl[L - 1][col], sxl[L][col], sxl[L + 1][col]};
    const T g0[2] = {sgx[L][col], sgy[L][col]};
    T gM[2] = {sgx[L - 1][col], sgy[L - 1][col]}, gP[2] = {sgx[L + 1][col], sgy[L + 1][col]};
    // a neighbour outside the domain has a zero gradient
    if (!inRow(L - 1)) gM[0] = gM[1] = T(0.0);
    if (!inRow(L + 1)) gP[0] = gP[1] = T(0.0);
    dualNodeFinish<R, false, T>(kp, a, t, na, nt, a * (nx + 1) + t, dpx, dpy, xls, g0, gM, gP);
  }
}

template <class R, int ROWS>
__global__ __launch_bounds__(kDTThreads) void k_dual_tile(KPT<R> kp) {
  // neighbouring column groups share the cache lines of the face fluxes
  // left of their first column: one XCD (one L2) takes a contiguous range
  dualTileBody<R, ROWS>(kp, xcdBlock());
}

unsigned dualTileBlock() { return kDTThreads; }

// the grid of tiles of `rows` node rows
dim3 dualTileGrid(long nx, long ny, int rows) {
  const long out = rows - 6;
  const long gx = (nx + 1 + kDTCols - 1) / kDTCols, gy = (ny + 1 + out - 1) / out;
  return dim3((unsigned)(gx * gy));
}

template <class R, bool XD>
__global__ __launch_bounds__(XD ? 256 : 64 * kDYWaves<R>) void k_dual_wave(KPT<R> kp) {
  // along y, neighbouring column groups share the cache lines of the face
  // fluxes left of their first column: one XCD (one L2) takes a contiguous
  // range of blocks
  dualWaveBody<R, XD>(kp, XD ? blockIdx.x : xcdBlock());
}
template <class R, bool XD, bool ONE>
__global__ __launch_bounds__(XD ? 256 : 64 * kDYWaves<R>) void k_dual_wave_batch(BatchLaunchT<R> bl) {
  batchRun<ONE>(bl, [&](const KPT<R>& kp, unsigned local, unsigned, unsigned) { dualWaveBody<R, XD>(kp, local); });
}

// threads per block of the wave dual kernels: 4 waves along x; along y 4
// (FP64) or 8 (mixed form)
unsigned dualWaveBlock(bool xd, bool mixed) {
  return xd ? 256u : 64u * (mixed ? kDYWaves<float> : kDYWaves<double>);
}

// launch grid of the wave dual kernels (blocks of dualWaveBlock threads)
dim3 dualWaveGrid(bool xd, long nx, long ny, bool mixed) {
  if (xd) {
    const long strips = (nx + 1 + kDXOut - 1) / kDXOut * (ny + 1);
    return dim3((unsigned)((strips + 3) / 4));
  }
  const int W = mixed ? kDYWaves<float> : kDYWaves<double>;
  const long gx = (nx + 1 + W * kDYCols - 1) / (W * kDYCols), gy = (ny + 1 + kDYOut - 1) / kDYOut;
  return dim3((unsigned)(gx * gy));
}

// Vnr::remapVariables, cells (PhaseRemapFinal.cc:23-205), then computeEOS
// and computePressionMoyenne (PhaseLagrange.cc:392-485) on the remapped state.
template <class R, int M>
__global__ __launch_bounds__(256) void k_remap_final(KPT<R> kp) {
  constexpr int NC = 3 * M + 1;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= kp.ncell) return;
  double U[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) U[q] = at(kp.f.U2[q], c);
  finalCell<R, M>(kp, c, U);
}

// computeNodeMass for output (PhaseLagrange.cc:31-48)
template <class R>
__global__ void k_node_mass(KPT<R> kp) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kp.nnode) return;
  const long nx = kp.o.nx, ny = kp.o.ny;
  const long j = p / (nx + 1), i = p - j * (nx + 1);
  const bool hc[4] = {j < ny && i < nx, j < ny && i > 0, j > 0 && i < nx, j > 0 && i > 0};
  const long cc[4] = {j * nx + i, j * nx + i - 1, (j - 1) * nx + i, (j - 1) * nx + i - 1};
  double s = 0.0;
  int n = 0;
  for (int q = 0; q < 4; ++q)
    if (hc[q]) s = s + at(kp.f.M, cc[q]), ++n;
  at(kp.f.Mp, p) = s / (double)n;
}

// Conservation sums with Neumaier compensation: out[0] mass, out[1]
// internal energy (sum M*e), out[2+k] material masses.  Per-block partial
// pairs (sum, compensation) are combined on the host.
// compact1: writes the derived material fields (fraction and mass fraction
// 1, the cell's mass, density, energy, pressure and sound speed) of the
// cells whose envx flag is clear into the material arrays of the current
// parity, so that every cell's material fields are readable there; the
// flags are left as they are.
template <class R>
__global__ __launch_bounds__(256) void k_env_expand(KPT<R> kp) {
  const auto& f = kp.f;
  const int s = kp.cur;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < kp.ncell; c += (long)gridDim.x * blockDim.x) {
    if (at(f.envx, c)) continue;
    at(f.f_k[0], c) = 1.0;
    at(f.mu_k[0], c) = R(1.0);
    at(f.M_k[0], c) = at(f.M, c);
    at(f.rho_k[s][0], c) = at(f.rho[s], c);
    at(f.e_k[s][0], c) = at(f.e[s], c);
    at(f.p_k[s][0], c) = at(f.p[s], c);
    at(f.c_k[s][0], c) = at(f.c[s], c);
  }
}

template <class R>
__global__ __launch_bounds__(256) void k_conservation(KPT<R> kp, double* out) {
  const auto& f = kp.f;
  const int m = kp.o.nbmat;
  const int nq = 2 + m;
  double s[2 + NMAT], comp[2 + NMAT];
  for (int q = 0; q < nq; ++q) s[q] = comp[q] = 0.0;
  auto add = [&](int q, double v) {
    const double t = s[q] + v;
    if (fabs(s[q]) >= fabs(v)) comp[q] += (s[q] - t) + v;
    else comp[q] += (v - t) + s[q];
    s[q] = t;
  };
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < kp.ncell; c += (long)gridDim.x * blockDim.x) {
    const double M = at(f.M, c);
    add(0, M);
    add(1, M * (double)at(f.e[kp.cur], c));
    // (compact1: the material mass is the cell mass unless stored)
    const bool derived = kp.o.compact1 && !at(f.envx, c);
    for (int k = 0; k < m; ++k) add(2 + k, derived ? M : at(f.M_k[k], c));
  }
  __shared__ double sh[256];
  for (int q = 0; q < nq; ++q) {
    for (int pass = 0; pass < 2; ++pass) {
      sh[threadIdx.x] = pass == 0 ? s[q] : comp[q];
      __syncthreads();
      if (threadIdx.x == 0) {
        double t = 0.0, cc = 0.0;
        for (int k = 0; k < (int)blockDim.x; ++k) {
          const double v = sh[k];
          const double y = t + v;
          if (fabs(t) >= fabs(v)) cc += (t - y) + v;
          else cc += (v - y) + t;
          t = y;
        }
        out[(blockIdx.x * nq + q) * 4 + pass * 2] = t;
        out[(blockIdx.x * nq + q) * 4 + pass * 2 + 1] = cc;
      }
      __syncthreads();
    }
  }
}

#define MHY_REMAP_INST(R)                                              \
  template __global__ void k_dual<R, true>(KPT<R>);                    \
  template __global__ void k_dual<R, false>(KPT<R>);                   \
  template __global__ void k_dual_batch<R, true, false>(BatchLaunchT<R>);  \
  template __global__ void k_dual_batch<R, false, false>(BatchLaunchT<R>); \
  template __global__ void k_dual_batch<R, true, true>(BatchLaunchT<R>);   \
  template __global__ void k_dual_batch<R, false, true>(BatchLaunchT<R>);  \
  template __global__ void k_dual_wave<R, true>(KPT<R>);               \
  template __global__ void k_dual_wave<R, false>(KPT<R>);              \
  template __global__ void k_dual_walk<R>(KPT<R>);                     \
  template __global__ void k_dual_wave_batch<R, true, false>(BatchLaunchT<R>);  \
  template __global__ void k_dual_wave_batch<R, false, false>(BatchLaunchT<R>); \
  template __global__ void k_dual_wave_batch<R, true, true>(BatchLaunchT<R>);   \
  template __global__ void k_dual_wave_batch<R, false, true>(BatchLaunchT<R>);  \
  template __global__ void k_remap_final<R, 1>(KPT<R>);                \
  template __global__ void k_remap_final<R, 2>(KPT<R>);                \
  template __global__ void k_remap_final<R, 3>(KPT<R>);                \
  template __global__ void k_node_mass<R>(KPT<R>);                     \
  template __global__ void k_conservation<R>(KPT<R>, double*);         \
  template __global__ void k_env_expand<R>(KPT<R>);                     \
  template __global__ void k_dual_tile<R, 32>(KPT<R>);              \
  template __global__ void k_dual_tile<R, 48>(KPT<R>);              \
  template __global__ void k_dual_tile<R, 64>(KPT<R>);
MHY_REMAP_INST(double)
MHY_REMAP_INST(float)
#undef MHY_REMAP_INST

}  // namespace mhy
