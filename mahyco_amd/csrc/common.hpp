// Shared host/device definitions of the VNR Lagrange + remap hot path:
// floating-point helpers with the exact semantics the reference relies on,
// the device state layout and the per-launch parameter block.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <cstring>

#include "../../include/mahyco.h"

#define MHY_HD __host__ __device__ __forceinline__

namespace mhy {

constexpr int NMAT = MHY_NBMATMAX;
// Live components of a cell remap vector: volume, mass, energy per material
// (3m), then the scheme's tail.  VNR: the pseudo-viscosity content
// (reference vector index 3m+3; indices 3m..3m+2 carry momentum and kinetic
// energy, always zero in VNR cell vectors).  Eucclhyd: the cell momentum
// (3m, 3m+1), and the kinetic energy (3m+2) when the projection is
// conservative; its pseudo-viscosity component is always zero.
constexpr int NCMAX = 3 * NMAT + 3;

// Layout of the remap vector, passed to the sweeps as their material
// template parameter L: L = m (1..3) is the VNR vector; L = 4 + m the
// Eucclhyd vector; L = 8 + m the Eucclhyd vector of a conservative
// projection.
template <int L>
constexpr int kMat = L & 3;
template <int L>
constexpr bool kEuc = L > 3;
template <int L>
constexpr int kTail = L < 4 ? 1 : (L < 8 ? 2 : 3);
template <int L>
constexpr int kNC = 3 * kMat<L> + kTail<L>;

// std::min / std::max argument conventions (cppnabla/types/MathFunctions.h:23-24)
MHY_HD double smin(double a, double b) { return (b < a) ? b : a; }
MHY_HD double smax(double a, double b) { return (a < b) ? b : a; }
// includes/Freefunctions.h:86-92
MHY_HD double divideNoExcept(double a, double b) { return (fabs(b) < 1.0E-12) ? 0.0 : a / b; }

struct D2 {
  double x, y;
};
MHY_HD D2 d2add(D2 a, D2 b) { return {a.x + b.x, a.y + b.y}; }
MHY_HD D2 d2sub(D2 a, D2 b) { return {a.x - b.x, a.y - b.y}; }
MHY_HD D2 d2mul(D2 a, double s) { return {a.x * s, a.y * s}; }
MHY_HD D2 d2div(D2 a, double s) { return {a.x / s, a.y / s}; }
// dot product accumulated from 0.0 (includes/Freefunctions.h:13-23)
MHY_HD double d2dot(D2 a, D2 b) {
  double r = 0.0;
  r = r + a.x * b.x;
  r = r + a.y * b.y;
  return r;
}
MHY_HD double d2norm(D2 a) { return sqrt(d2dot(a, a)); }
MHY_HD double d2cross(D2 a, D2 b) { return a.x * b.y - a.y * b.x; }

// corner normal |C|.(C/|C|) per half edge (includes/Freefunctions.h:25-58)
MHY_HD D2 lpcNpc(D2 xp, D2 xpPlus, D2 xpMinus) {
  D2 n1 = {0.5 * (xpPlus.y - xp.y), 0.5 * (xp.x - xpPlus.x)};
  double l1 = d2norm(n1);
  D2 n2 = {0.5 * (xp.y - xpMinus.y), 0.5 * (xpMinus.x - xp.x)};
  double l2 = d2norm(n2);
  return d2add(d2mul(d2div(n1, l1), l1), d2mul(d2div(n2, l2), l2));
}
// The same corner normal on the Euler grid.  Every half edge there is
// axis-aligned, so each half-edge vector has one zero component and
// |n| = sqrt(fl(a*a)) = |a| exactly (binary floating point), n/|n| = +-1 and
// (n/|n|)*|n| = n: the sum of the two raw half-edge vectors is bitwise the
// value of lpcNpc, without its 2 square roots and 4 divisions.
MHY_HD D2 lpcNpcAxis(D2 xp, D2 xpPlus, D2 xpMinus) {
  return {0.5 * (xpPlus.y - xp.y) + 0.5 * (xp.y - xpMinus.y), 0.5 * (xp.x - xpPlus.x) + 0.5 * (xpMinus.x - xp.x)};
}
// includes/Freefunctions.h:60-65
MHY_HD D2 symmetricVector(D2 v, D2 sigma) { return d2sub(d2mul(sigma, 2.0 * d2dot(v, sigma)), v); }

// sub-volume of corner s of a quad (vnr_remap/PhaseLagrange.cc:236-264)
MHY_HD double subVolume(D2 xc, D2 xm, D2 x0, D2 xp) {
  const D2 x2 = d2mul(d2add(xm, x0), 0.5);
  const D2 x4 = d2mul(d2add(xp, x0), 0.5);
  return 0.5 * (d2cross(xc, x2) + d2cross(x2, x0) + d2cross(x0, x4) + d2cross(x4, xc));
}

// Run mode of a context.
enum Mode : int {
  MODE_FULL = 0,       // Lagrange + remap (Euler grid fixed)
  MODE_LAGRANGE = 1,   // AvecProjection = 0: moving mesh
  MODE_ADVECTION = 2,  // sansLagrange = 1: remap of an imposed velocity field
};

// Time-loop scalars, double-buffered like the reference's n/n+1 pairs.
// minKey / maxKey accumulate the CFL minimum and the maximum pressure of the
// step with parity cur as order-preserving integer keys (atomic min / max,
// exact and order-free); the finalize kernel turns them into dt and pmax
// and resets the other parity's pair for the next step.
struct Scalars {
  double t[2];
  double dt[2];
  double pmax;
  int err;        // MHY_ERR_* raised by a kernel (energy abort, NaN)
  int pad;
  unsigned long long minKey[2], maxKey[2];
};

// Integer keys ordered like the doubles they encode.  minKey only sees
// values >= +0 or NaN (CFL ratios), whose bit patterns already order as
// unsigned integers, with every NaN above the initial DBL_MAX; maxKey maps
// the sign so that all doubles order, and NaN is filtered before.
MHY_HD unsigned long long dkeyPos(double v) { return __builtin_bit_cast(unsigned long long, v); }
MHY_HD double dkeyPosInv(unsigned long long k) { return __builtin_bit_cast(double, k); }
MHY_HD unsigned long long dkeySigned(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
MHY_HD double dkeySignedInv(unsigned long long k) {
  const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __builtin_bit_cast(double, b);
}

// Options and geometry constants, uniform for a launch.
struct Opts {
  int nx, ny, nbmat, nc;  // nc live remap components (kNC of the layout)
  int mode;
  int euc;                // Eucclhyd scheme (cell-centred velocities)
  int space_order;        // Eucclhyd: spaceOrder (1 or 2)
  int conservative;       // Eucclhyd: projectionConservative
  double deltat_min;
  int order;              // projectionOrder
  int limiter, limiter_pure, pp, pp_mixte, pp_complet;
  int pseudo_centree, methode, abort_energy;
  int eos[NMAT];
  double gammap[NMAT];
  double gamma, threshold, dx, dy, cfl, cfl_euler;
  // boundary conditions (includes/ConditionsLimites.h:9-43)
  int bc_left, bc_right, bc_bottom, bc_top;
  double bcv_left[2], bcv_right[2], bcv_bottom[2], bcv_top[2];
  int periodic_right;
  int flux_left, flux_right, flux_bottom, flux_top, flux_any;
  // one material, perfect gas, Lagrange + remap: the material fields
  // (fraction, mass fraction, mass, density, energy, pressure, sound speed)
  // are stored only for cells whose envx flag is set; elsewhere they equal
  // 1, 1 and the cell's mean fields bit for bit
  int compact1;
};

// Device pointers of one problem (structure of arrays, cell index j*nx+i,
// node index j*(nx+1)+i).  [2] arrays are the reference's n/n+1 pairs,
// addressed through the parity of the launch.  R is the storage type of the
// bulk cell / node fields: double for the FP64 form, float for the mixed
// form.  Volume fractions, masses, the remap vectors and fluxes, the nodal
// momentum of the dual remap, coordinates and the time-step scalars are FP64
// in both forms.
template <class R>
struct FieldsT {
  R *rho[2], *p[2], *q[2], *c[2], *divu[2], *e[2];
  R *rho_k[2][NMAT], *p_k[2][NMAT], *q_k[2][NMAT], *c_k[2][NMAT], *e_k[2][NMAT];
  R *ux[2], *uy[2];
  double *X[2], *Y[2];  // moving node coordinates (MODE_LAGRANGE)
  double *vol[2];       // sum of the corner sub-volumes of the step start
  double *f_k[NMAT];
  R *mu_k[NMAT];
  double *M, *M_k[NMAT];
  int *mixte, *pure;
  R *Mp, *outvx, *outvy;
  R *rho0, *e0, *rho_k0[NMAT], *e_k0[NMAT];  // t=0 fields (periodic cells)
  // per-step intermediates
  double *UL[NCMAX], *U1[NCMAX], *U2[NCMAX];
  double *VL, *XcL, *YcL, *rL, *qL;
  double *G[NMAT];   // face fluxes of the current sweep for the dual remap (A1/A2)
  double *DF[NMAT];  // PB dual flux per cell of the current sweep
  double *UD1[3];    // nodal vector after dual sweep 1 (M ux, M uy, M)
  double *XL2;       // end-of-Lagrange node coordinate along the second sweep
  unsigned char* envx;  // compact1: the cell's material fields are stored (1) or derived (0)
  Scalars* sc;
};

// Parameters of one launch: fields, options, parity and sweep direction.
template <class R>
struct KPT {
  FieldsT<R> f;
  Opts o;
  int cur, nxt;  // parity of the n and n+1 buffers
  int sweep;     // 1 or 2
  int xdir;      // 1: sweep along x, 0: along y
  int fuseFinal; // sweep 2 finishes each cell (final remap) instead of storing U2
  int geomLoad;  // the sweeps read the Lagrange volume and centroid (V_L, Xc_L, Yc_L)
                 // instead of forming them again from the node velocities
  long ncell, nnode;
};

using Fields = FieldsT<double>;
using KP = KPT<double>;

// Arrays of the Eucclhyd scheme (FP64), beside the shared fields: the cell
// velocity, the velocity gradient, the sum over each cell's corners of the
// material corner forces of the last step (the next step's material
// pressure gradients, times the cell volume), the nodes' ranges of the
// material pressures and of the cell velocity (the second-order
// extrapolation), the cell centroid and volume ([2]: they move with the mesh
// without remap) and perimeter, and the output velocities.
struct EucFields {
  double *vx[2], *vy[2];
  double *gv[4];
  double *fsx[NMAT], *fsy[NMAT];
  double *pmin[NMAT], *pmax[NMAT];
  double *vmin[2], *vmax[2];
  double *xc[2], *yc[2], *ve[2];
  double *perim;
  double *outvx, *outvy;
};

// Element idx of a device array addressed by a 32-bit byte offset: the load
// or store takes the array base from scalar registers and the offset from one
// 32-bit vector register, instead of a 64-bit address per array.  Every array
// of a context is below 4 GiB (checked when the context is created).
template <class V>
__device__ __forceinline__ V& at(V* base, long idx) {
  return *reinterpret_cast<V*>(reinterpret_cast<char*>(base) + (unsigned)idx * (unsigned)sizeof(V));
}
template <class V>
__device__ __forceinline__ const V& at(const V* base, long idx) {
  return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + (unsigned)idx * (unsigned)sizeof(V));
}

// bitwise equality (distinguishes signed zeros, never true for a NaN pattern
// compared with another)
__device__ __forceinline__ bool sameBits(double a, double b) {
  return __double_as_longlong(a) == __double_as_longlong(b);
}
__device__ __forceinline__ bool sameBits(float a, float b) { return __float_as_int(a) == __float_as_int(b); }

// true for the mixed form (FP32 bulk storage)
template <class R>
constexpr bool kMixed = sizeof(R) == 4;

// The same parameter block seen with another bulk storage type: every field
// member is a pointer, so the layouts are identical and the host keeps one
// untyped copy (KP) whatever the form.
template <class R>
KPT<R> retype(const KP& kp) {
  static_assert(sizeof(KPT<R>) == sizeof(KP), "parameter blocks differ");
  KPT<R> r;
  memcpy(&r, &kp, sizeof(r));
  return r;
}

// Logical block of this workgroup.  The hardware hands consecutive
// workgroup ids to the 8 XCDs in turn, and each XCD has its own L2; the
// renumbering gives every XCD one contiguous range of logical blocks (a band
// of the mesh), so that the cache lines two neighbouring blocks share are
// fetched into one L2 only.  It is a permutation of [0, gridDim.x).  Used by
// the y sweeps, whose blocks cover 16-column groups (64 bytes of an FP32
// row); measured slower for the kernels that stream whole rows.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcdBlock(unsigned b, unsigned nblocks) {
  const unsigned q = nblocks / kXcds, r = nblocks % kXcds;
  const unsigned k = b % kXcds, i = b / kXcds;
  return k * q + (k < r ? k : r) + i;
}
__device__ __forceinline__ unsigned xcdBlock() { return xcdBlock(blockIdx.x, gridDim.x); }

// DPP lane moves (a VALU operand modifier, no LDS round trip): wave_shr:1
// gives each lane the value of lane - 1, wave_shl:1 that of lane + 1; the
// lane with no source (0, resp. 63) keeps its own value, as a shuffle clamped
// to the lane itself does.  Called in converged control flow only.
constexpr int kDppWaveShr1 = 0x138, kDppWaveShl1 = 0x130;
template <int CTRL>
__device__ __forceinline__ int dppMove(int v) {
  return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ float dppMove(float v) {
  return __int_as_float(dppMove<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dppMove(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dppMove<CTRL>((int)(b & 0xffffffffLL)), hi = dppMove<CTRL>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// One launch over an ensemble of problems: the launch parameters of problem
// p for step variant v are table[v * nprob + p]; the virtual blocks of problem
// p are start[p] .. start[p+1]-1, laid out as rows of gx[p] blocks, and
// owner[vb] is the problem of virtual block vb.
template <class R>
struct BatchLaunchT {
  const KPT<R>* table;
  const unsigned* start;  // nprob + 1 entries
  const unsigned* gx;     // nprob entries
  int nprob, variant;
  unsigned total;         // start[nprob]
  const unsigned* owner;  // total entries
};
using BatchLaunch = BatchLaunchT<double>;

// problem p and block index inside it of virtual block vb
template <class R>
__device__ __forceinline__ void batchLocate(const BatchLaunchT<R>& bl, unsigned vb, int& p, unsigned& local) {
  p = (int)bl.owner[vb];
  local = vb - bl.start[p];
}

// Runs fn(kp, local block, gx, blocks of the problem) for a contiguous range
// of virtual blocks: workgroup w takes blocks [w*per, (w+1)*per), so that it
// stays on one problem (and its launch parameters) for as long as possible.
// The first problem comes from the owner table (one load); the walk then
// moves on through start[].
template <class R, class F>
__device__ __forceinline__ void batchForEach(const BatchLaunchT<R>& bl, F&& fn) {
  const unsigned per = (bl.total + gridDim.x - 1) / gridDim.x;
  unsigned vb = blockIdx.x * per;
  const unsigned end = (vb + per < bl.total) ? vb + per : bl.total;
  if (vb >= end) return;
  int p;
  unsigned local;
  batchLocate(bl, vb, p, local);
  unsigned pend = bl.start[p + 1];
  for (; vb < end; ++vb, ++local) {
    while (vb >= pend) {  // next problem (problems without blocks are skipped)
      ++p;
      local = 0;
      pend = bl.start[p + 1];
    }
    fn(bl.table[bl.variant * bl.nprob + p], local, bl.gx[p], pend - bl.start[p]);
  }
}

// One workgroup per virtual block (ONE, a launch of start[nprob] workgroups)
// or the contiguous ranges of batchForEach.  Separate kernels: the walk's
// state costs registers (occupancy) that a single block does not need.
template <bool ONE, class R, class F>
__device__ __forceinline__ void batchRun(const BatchLaunchT<R>& bl, F&& fn) {
  if (ONE) {
    const unsigned vb = blockIdx.x;
    if (vb >= bl.total) return;
    const int p = (int)bl.owner[vb];
    const unsigned s0 = bl.start[p];
    fn(bl.table[bl.variant * bl.nprob + p], vb - s0, bl.gx[p], bl.start[p + 1] - s0);
  } else {
    batchForEach(bl, fn);
  }
}

// Euler node coordinates (CartesianMesh2DGenerator.cc:52-55: x = dx*i, y = dy*j)
MHY_HD D2 eulerNode(const Opts& o, long j, long i) { return {o.dx * (double)i, o.dy * (double)j}; }

}  // namespace mhy
