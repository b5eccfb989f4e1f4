// Lagrange phase of the VNR scheme on CDNA4: CFL time-step / max-pressure
// reduction, nodal momentum update with boundary conditions, and the cell
// update (sub-volumes, densities, artificial viscosity, energy, perfect-gas
// EOS) fused with the preparation of the remap vectors.
//
// Arithmetic follows the reference operation by operation (compiled with
// -ffp-contract=off), so FP64 results are bitwise those of the CPU path.
#include "common.hpp"
#include "kernels.hpp"
#include "lagrange_cell.hpp"

namespace mhy {

namespace {

__device__ __forceinline__ double waveMin(double v) {
  for (int o = 32; o > 0; o >>= 1) v = smin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double waveMax(double v) {
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_xor(v, o, 64);
    v = (w > v) ? w : v;
  }
  return v;
}

// node coordinates at the step start
template <class R>
__device__ __forceinline__ D2 nodeX(const KPT<R>& kp, long j, long i) {
  if (kp.o.mode == MODE_LAGRANGE) {
    long p = j * (kp.o.nx + 1) + i;
    return {kp.f.X[kp.cur][p], kp.f.Y[kp.cur][p]};
  }
  return eulerNode(kp.o, j, i);
}

}  // namespace

// vnr_remap/VnrRemap.cc:14-50 (computeDeltaT) fused with the max-pressure
// reduction of PhaseLagrange.cc:55-61.  Min and max ignore NaN values like
// the reference's reducers.  Each block reduces its cells and adds them to
// the step's accumulators with one 64-bit atomic min / max (exact and
// order-free, no fence); k_dt_finalize then forms dt and t (computeTime).
// b: block index among the nb blocks working on this problem.  The
// reduction is FP64 in both forms (FP32 inputs are widened).
template <class R>
__device__ __forceinline__ void dtReduceBody(const KPT<R>& kp, unsigned b, unsigned nb) {
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const long nx = o.nx;
  double vmin = DBL_MAX, pm = -DBL_MAX;
  const R* ux = f.ux[kp.cur];
  const R* uy = f.uy[kp.cur];
  for (long c = (long)b * blockDim.x + threadIdx.x; c < kp.ncell; c += (long)nb * blockDim.x) {
    // 32-bit division: cell and node counts stay below 2^31 (checked at set-up)
    const long j = (long)((unsigned)c / (unsigned)nx), i = c - j * nx;
    const long p0 = j * (nx + 1) + i;
    const long nd[4] = {p0, p0 + 1, p0 + nx + 2, p0 + nx + 1};
    double uc = 0.0;
    for (int s = 0; s < 4; ++s) {
      const double a = ux[nd[s]], b = uy[nd[s]];
      uc += sqrt(a * a + b * b) * 0.25;
    }
    const double v = o.cfl * sqrt(f.vol[kp.cur][c]) / (o.cfl_euler * uc + (double)f.c[kp.cur][c]);
    vmin = smin(vmin, v);
    const double pc = f.p[kp.cur][c];
    pm = (pc > pm) ? pc : pm;
  }
  __shared__ double smn[4], smx[4];
  vmin = waveMin(vmin);
  pm = waveMax(pm);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) smn[w] = vmin, smx[w] = pm;
  __syncthreads();
  if (threadIdx.x == 0) {
    double lo = smn[0], hi = smx[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) lo = smin(lo, smn[k]), hi = (smx[k] > hi) ? smx[k] : hi;
    Scalars* sc = f.sc;
    atomicMin(&sc->minKey[kp.cur], dkeyPos(lo));
    atomicMax(&sc->maxKey[kp.cur], dkeySigned(hi));
  }
}

template <class R>
__global__ __launch_bounds__(256) void k_dt_reduce(KPT<R> kp) {
  dtReduceBody(kp, blockIdx.x, gridDim.x);
}

// dt_{n+1} = min(CFL minimum, 1.05 dt_n) and t_{n+1} (VnrRemap.cc:45-50,
// computeTime); resets the accumulators of the next step's parity.
__device__ __forceinline__ void dtFinalize(Scalars* sc, int cur, int nxt) {
  const double lo = dkeyPosInv(sc->minKey[cur]);
  const double dtn = sc->dt[cur];
  const double dt1 = smin(lo, 1.05 * dtn);
  sc->dt[nxt] = dt1;
  sc->t[nxt] = sc->t[cur] + dt1;
  sc->pmax = dkeySignedInv(sc->maxKey[cur]);
  sc->minKey[nxt] = dkeyPos(DBL_MAX);
  sc->maxKey[nxt] = dkeySigned(-DBL_MAX);
}

template <class R>
__global__ void k_dt_finalize(KPT<R> kp) {
  if (threadIdx.x == 0 && blockIdx.x == 0) dtFinalize(kp.f.sc, kp.cur, kp.nxt);
}

// sansLagrange: deltat_nplus1 = deltat_n (VnrRemap.cc:143-146)
template <class R>
__global__ void k_dt_hold(KPT<R> kp) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    Scalars* sc = kp.f.sc;
    sc->dt[kp.nxt] = sc->dt[kp.cur];
    sc->t[kp.nxt] = sc->t[kp.cur] + sc->dt[kp.nxt];
  }
}

// Node kernel: node mass as the mean of the adjacent cell masses
// (PhaseLagrange.cc:31-48), then updateVelocity (PhaseLagrange.cc:167-197),
// updateVelocityBoundaryConditions (ConditionsLimites.cc:12-185) and
// updatePosition (PhaseLagrange.cc:203-208).  In advection mode the velocity
// is carried over unchanged (VnrRemap.cc:152-154).  Arithmetic is FP64 in
// both forms.
template <class R>
__device__ __forceinline__ void lagrangeNodeAt(const KPT<R>& kp, long p, long j, long i) {
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const long nx = o.nx, ny = o.ny;
  // cells of the node in getCellsOfNode order (CartesianMesh2D.cc:97-108)
  const bool hc[4] = {j < ny && i < nx, j < ny && i > 0, j > 0 && i < nx, j > 0 && i > 0};
  const long cc[4] = {j * nx + i, j * nx + i - 1, (j - 1) * nx + i, (j - 1) * nx + i - 1};
  double msum = 0.0;
  int cnt = 0;
  for (int q = 0; q < 4; ++q)
    if (hc[q]) msum = msum + f.M[cc[q]], ++cnt;
  const double Mp = msum / (double)cnt;
  f.Mp[p] = Mp;
  const double uxn = (double)f.ux[kp.cur][p], uyn = (double)f.uy[kp.cur][p];
  if (o.mode == MODE_ADVECTION) {
    f.ux[kp.nxt][p] = uxn;
    f.uy[kp.nxt][p] = uyn;
    return;
  }
  const Scalars* sc = f.sc;
  const double dt1 = sc->dt[kp.nxt], dtn = sc->dt[kp.cur];
  const double dt = 0.5 * (dt1 + dtn);
  // corner normals of the four cells at this node: C(cell, node) depends on
  // the node and its two neighbours along the cell's edges
  const D2 X0 = nodeX(kp, j, i);
  D2 Cn[4];
  if (o.mode == MODE_LAGRANGE) {
    if (hc[0]) Cn[0] = lpcNpc(X0, nodeX(kp, j, i + 1), nodeX(kp, j + 1, i));
    if (hc[1]) Cn[1] = lpcNpc(X0, nodeX(kp, j + 1, i), nodeX(kp, j, i - 1));
    if (hc[2]) Cn[2] = lpcNpc(X0, nodeX(kp, j - 1, i), nodeX(kp, j, i + 1));
    if (hc[3]) Cn[3] = lpcNpc(X0, nodeX(kp, j, i - 1), nodeX(kp, j - 1, i));
  } else {  // fixed Euler grid: axis-aligned edges
    if (hc[0]) Cn[0] = lpcNpcAxis(X0, eulerNode(o, j, i + 1), eulerNode(o, j + 1, i));
    if (hc[1]) Cn[1] = lpcNpcAxis(X0, eulerNode(o, j + 1, i), eulerNode(o, j, i - 1));
    if (hc[2]) Cn[2] = lpcNpcAxis(X0, eulerNode(o, j - 1, i), eulerNode(o, j, i + 1));
    if (hc[3]) Cn[3] = lpcNpcAxis(X0, eulerNode(o, j, i - 1), eulerNode(o, j - 1, i));
  }
  double pq[4];
  for (int q = 0; q < 4; ++q)
    if (hc[q]) pq[q] = (double)f.p[kp.cur][cc[q]] + (double)f.q[kp.cur][cc[q]];
  const bool inner = (i != 0 && j != 0 && i != nx && j != ny);
  D2 un = {uxn, uyn};
  D2 u1;
  if (inner) {
    D2 r = {0.0, 0.0};
    for (int q = 0; q < 4; ++q) r = d2add(r, d2mul(Cn[q], pq[q]));
    u1 = d2add(un, d2mul(r, dt / Mp));
    if (o.mode == MODE_LAGRANGE) f.outvx[p] = u1.x, f.outvy[p] = u1.y;
  } else {
    // sides in the reference's launch order, then corners; a side that is
    // neither symmetry nor imposed velocity leaves the buffer unchanged
    u1 = {(double)f.ux[kp.nxt][p], (double)f.uy[kp.nxt][p]};
    auto side = [&](int bc, D2 axis, const double* val) {
      if (bc == MHY_BC_SYMMETRY) {
        D2 r = {0.0, 0.0};
        for (int q = 0; q < 4; ++q)
          if (hc[q]) r = d2add(r, d2mul(d2add(Cn[q], symmetricVector(Cn[q], axis)), pq[q]));
        u1 = d2add(un, d2mul(r, dt / (Mp)));
      } else if (bc == MHY_BC_IMPOSED_VELOCITY) {
        u1 = {val[0], val[1]};
      }
    };
    const D2 ax = {1.0, 0.0}, ay = {0.0, 1.0};
    if (j == 0) side(o.bc_bottom, ax, o.bcv_bottom);
    if (j == ny) side(o.bc_top, ax, o.bcv_top);
    if (i == 0) side(o.bc_left, ay, o.bcv_left);
    if (i == nx) side(o.bc_right, ay, o.bcv_right);
    auto corner = [&](int a, int b, const double* val) {
      if (a == MHY_BC_SYMMETRY && b == MHY_BC_SYMMETRY) u1 = {0.0, 0.0};
      else if (a == MHY_BC_IMPOSED_VELOCITY && b == MHY_BC_IMPOSED_VELOCITY) u1 = {val[0], val[1]};
    };
    if (j == ny && i == 0) corner(o.bc_top, o.bc_left, o.bcv_left);
    if (j == ny && i == nx) corner(o.bc_top, o.bc_right, o.bcv_right);
    if (j == 0 && i == 0) corner(o.bc_bottom, o.bc_left, o.bcv_left);
    if (j == 0 && i == nx) corner(o.bc_bottom, o.bc_right, o.bcv_right);
  }
  f.ux[kp.nxt][p] = u1.x;
  f.uy[kp.nxt][p] = u1.y;
  if (o.mode == MODE_LAGRANGE) {
    f.X[kp.nxt][p] = X0.x + u1.x * dt1;
    f.Y[kp.nxt][p] = X0.y + u1.y * dt1;
  }
}

template <class R>
__device__ __forceinline__ void lagrangeNodeBody(const KPT<R>& kp, unsigned b) {
  const long p = (long)b * blockDim.x + threadIdx.x;
  if (p >= kp.nnode) return;
  const long nx = kp.o.nx;
  const long j = (long)((unsigned)p / (unsigned)(nx + 1)), i = p - j * (nx + 1);
  lagrangeNodeAt(kp, p, j, i);
}

template <class R>
__global__ __launch_bounds__(256) void k_lagrange_node(KPT<R> kp) {
  lagrangeNodeBody(kp, blockIdx.x);
}

// The same node update split by position, for the fixed Euler grid
// (MODE_FULL): the interior nodes, whose four cells all exist and whose
// corner normals are the raw half-edge sums, run in a kernel of their own
// (blockIdx.y = node row, no index division, few registers); the boundary
// nodes, with their conditions, run the general body above.  Each node is
// updated by exactly one of the two, from the same operands in the same
// order, so the results are bitwise those of k_lagrange_node.
template <class R>
__global__ __launch_bounds__(256) void k_lagrange_node_inner(KPT<R> kp) {
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const int nx = o.nx;
  const int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int j = 1 + (int)blockIdx.y;
  if (i >= nx) return;
  const int p = j * (nx + 1) + i;
  const int c0 = j * nx + i;  // cells (j, i), (j, i-1), (j-1, i), (j-1, i-1)
  const int cc[4] = {c0, c0 - 1, c0 - nx, c0 - nx - 1};
  double msum = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) msum = msum + f.M[cc[q]];
  const double Mp = msum / 4.0;
  f.Mp[p] = Mp;
  const double uxn = (double)f.ux[kp.cur][p], uyn = (double)f.uy[kp.cur][p];
  const Scalars* sc = f.sc;
  const double dt1 = sc->dt[kp.nxt], dtn = sc->dt[kp.cur];
  const double dt = 0.5 * (dt1 + dtn);
  const D2 X0 = eulerNode(o, j, i);
  const D2 Cn[4] = {lpcNpcAxis(X0, eulerNode(o, j, i + 1), eulerNode(o, j + 1, i)),
                    lpcNpcAxis(X0, eulerNode(o, j + 1, i), eulerNode(o, j, i - 1)),
                    lpcNpcAxis(X0, eulerNode(o, j - 1, i), eulerNode(o, j, i + 1)),
                    lpcNpcAxis(X0, eulerNode(o, j, i - 1), eulerNode(o, j - 1, i))};
  D2 r = {0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double pq = (double)f.p[kp.cur][cc[q]] + (double)f.q[kp.cur][cc[q]];
    r = d2add(r, d2mul(Cn[q], pq));
  }
  const D2 u1 = d2add(D2{uxn, uyn}, d2mul(r, dt / Mp));
  f.ux[kp.nxt][p] = u1.x;
  f.uy[kp.nxt][p] = u1.y;
}

// the boundary nodes: bottom row, top row, then the left and right columns
// without their corners
template <class R>
__global__ __launch_bounds__(256) void k_lagrange_node_edge(KPT<R> kp) {
  const long nx = kp.o.nx, ny = kp.o.ny;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long j, i;
  if (e < nx + 1) {
    j = 0, i = e;
  } else if (e < 2 * (nx + 1)) {
    j = ny, i = e - (nx + 1);
  } else if (e < 2 * (nx + 1) + (ny - 1)) {
    j = 1 + (e - 2 * (nx + 1)), i = 0;
  } else if (e < 2 * (nx + 1) + 2 * (ny - 1)) {
    j = 1 + (e - 2 * (nx + 1) - (ny - 1)), i = nx;
  } else {
    return;
  }
  lagrangeNodeAt(kp, j * (nx + 1) + i, j, i);
}

// Cell kernel: the cell part of the Lagrange phase (lagrangeCellCompute,
// lagrange_cell.hpp) with the remap vector, the Lagrange volume and centroid
// (and the Lagrange density / pseudo where the options need them) stored for
// the first remap sweep.  The wave sweeps fuse this update into their first
// sweep instead; this kernel serves the other sweep kernels and the
// moving-mesh mode.
template <class R, int M>
__device__ __forceinline__ void lagrangeCellBody(const KPT<R>& kp, unsigned b) {
  const long c = (long)b * blockDim.x + threadIdx.x;
  if (c >= kp.ncell) return;
  const long nx = kp.o.nx;
  const long j = (long)((unsigned)c / (unsigned)nx), i = c - j * nx;
  LagCell<M> lc;
  if (!lagrangeCellCompute<R, M>(kp, c, j, i, true, lc)) return;
  const auto& f = kp.f;
#pragma unroll
  for (int q = 0; q < 3 * M + 1; ++q) at(f.UL[q], c) = lc.U[q];
  at(f.VL, c) = lc.V;
  if (kp.o.methode == 1) at(f.rL, c) = lc.rho;
  if (kp.o.pp == 1) at(f.qL, c) = lc.q;
  at(f.XcL, c) = lc.xc.x;
  at(f.YcL, c) = lc.xc.y;
}

template <class R, int M>
__global__ __launch_bounds__(256) void k_lagrange_cell(KPT<R> kp) {
  lagrangeCellBody<R, M>(kp, blockIdx.x);
}

// ---------------------------------------------------------------- batches
// One launch serves every problem of an ensemble: the virtual blocks of all
// problems are numbered consecutively (BatchLaunch::start) and a fixed
// number of workgroups walks them.

template <class R, bool ONE>
__global__ __launch_bounds__(256) void k_dt_reduce_batch(BatchLaunchT<R> bl) {
  batchRun<ONE>(bl, [&](const KPT<R>& kp, unsigned b, unsigned, unsigned nb) {
    dtReduceBody(kp, b, nb);
    __syncthreads();  // the shared scratch is reused by the next virtual block
  });
}
// one thread per problem
template <class R>
__global__ __launch_bounds__(256) void k_dt_finalize_batch(BatchLaunchT<R> bl) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= bl.nprob) return;
  const KPT<R>& kp = bl.table[bl.variant * bl.nprob + p];
  dtFinalize(kp.f.sc, kp.cur, kp.nxt);
}
template <class R, bool ONE>
__global__ __launch_bounds__(256) void k_lagrange_node_batch(BatchLaunchT<R> bl) {
  batchRun<ONE>(bl, [&](const KPT<R>& kp, unsigned b, unsigned, unsigned) { lagrangeNodeBody(kp, b); });
}
template <class R, int M, bool ONE>
__global__ __launch_bounds__(256) void k_lagrange_cell_batch(BatchLaunchT<R> bl) {
  batchRun<ONE>(bl, [&](const KPT<R>& kp, unsigned b, unsigned, unsigned) { lagrangeCellBody<R, M>(kp, b); });
}

#define MHY_LAGRANGE_INST_ONE(R, ONE)                                        \
  template __global__ void k_dt_reduce_batch<R, ONE>(BatchLaunchT<R>);       \
  template __global__ void k_lagrange_node_batch<R, ONE>(BatchLaunchT<R>);   \
  template __global__ void k_lagrange_cell_batch<R, 1, ONE>(BatchLaunchT<R>); \
  template __global__ void k_lagrange_cell_batch<R, 2, ONE>(BatchLaunchT<R>); \
  template __global__ void k_lagrange_cell_batch<R, 3, ONE>(BatchLaunchT<R>);
#define MHY_LAGRANGE_INST(R)                                                 \
  template __global__ void k_dt_reduce<R>(KPT<R>);                           \
  template __global__ void k_dt_finalize<R>(KPT<R>);                         \
  template __global__ void k_dt_finalize_batch<R>(BatchLaunchT<R>);          \
  template __global__ void k_dt_hold<R>(KPT<R>);                             \
  template __global__ void k_lagrange_node<R>(KPT<R>);                       \
  template __global__ void k_lagrange_node_inner<R>(KPT<R>);                 \
  template __global__ void k_lagrange_node_edge<R>(KPT<R>);                  \
  template __global__ void k_lagrange_cell<R, 1>(KPT<R>);                    \
  template __global__ void k_lagrange_cell<R, 2>(KPT<R>);                    \
  template __global__ void k_lagrange_cell<R, 3>(KPT<R>);                    \
  MHY_LAGRANGE_INST_ONE(R, false)                                            \
  MHY_LAGRANGE_INST_ONE(R, true)
MHY_LAGRANGE_INST(double)
MHY_LAGRANGE_INST(float)
#undef MHY_LAGRANGE_INST
#undef MHY_LAGRANGE_INST_ONE

}  // namespace mhy
