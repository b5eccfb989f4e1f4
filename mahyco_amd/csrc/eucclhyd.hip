// The cell-centred Eucclhyd Lagrange phase (eucclhyd_remap/PhaseLagrange.cc,
// ConditionsLimites.cc, EucclhydRemap.cc) and its final remap
// (PhaseRemapFinal.cc), the second caller of the remap cell sweeps.  Three
// kernels per step, then the sweeps on the Eucclhyd remap-vector layout:
//
//  k_euc_start (cells): equation of state and mean pressure, velocity
//    gradient from the node velocities at n, cell masses, and the cell time
//    step into the step's minimum (computeEOS .. computem_cell_deltat);
//  k_euc_node (nodes): the node's ranges of the material pressures and cell
//    velocities, the corner extrapolations, dissipation matrices and G of its
//    cells, the node velocity (inner nodes and boundary conditions)
//    (extrapolateValue .. computeNodeVelocity);
//  k_euc_cell (cells): the corner forces from the same extrapolations (formed
//    again from the node ranges instead of stored per corner), the Lagrange
//    volume and centroid, density, velocity, material energies and the remap
//    vector U_L (computeSubCellForce .. updateCellCenteredLagrangeVariables);
//    without remap, the state n+1.
//
// The reference keeps corner quantities per node and cell of the node; here
// each is formed where it is used, from the Euler node coordinates (the
// corner normals of the set-up mesh) and the cell and node fields, with the
// reference's operation order.  FP64 throughout (-ffp-contract=off): results
// are bitwise those of the CPU path.
#include "final_math.hpp"
#include "kernels.hpp"

namespace mhy {
namespace {

struct M22 {
  double a, b, c, d;  // row-major
};
// types/MathFunctions.h:41-47: each row a dot product from 0.0
__device__ __forceinline__ D2 matVec(M22 m, D2 v) { return {d2dot({m.a, m.b}, v), d2dot({m.c, m.d}, v)}; }
// eucclhyd_remap/Utiles.cc:6-10
__device__ __forceinline__ M22 inverse2(M22 m) {
  const double alpha = 1.0 / (m.a * m.d - m.b * m.c);
  return {m.d * alpha, -m.b * alpha, -m.c * alpha, m.a * alpha};
}
// eucclhyd_remap/Utiles-Impl.h:4-12
__device__ __forceinline__ M22 tens(D2 a, D2 b) { return {a.x * b.x, a.x * b.y, a.y * b.x, a.y * b.y}; }
__device__ __forceinline__ M22 madd(M22 p, M22 q) { return {p.a + q.a, p.b + q.b, p.c + q.c, p.d + q.d}; }
__device__ __forceinline__ M22 mscale(double s, M22 m) { return {s * m.a, s * m.b, s * m.c, s * m.d}; }

constexpr double kEucCfl = 0.45;  // includes/GestionTemps.h:11 (never read from input)

// nodes of cell (j, i) in the quad order BL, BR, TR, TL
__device__ __forceinline__ void quadNodes(long nx, long j, long i, long* jj, long* ii, long* p) {
  jj[0] = j, ii[0] = i, jj[1] = j, ii[1] = i + 1, jj[2] = j + 1, ii[2] = i + 1, jj[3] = j + 1, ii[3] = i;
  for (int s = 0; s < 4; ++s) p[s] = jj[s] * (nx + 1) + ii[s];
}

// Corner s of cell (j, i) on the set-up mesh (computeCornerNormal,
// PhaseLagrange.cc:27-66): half-edge normals to the next and previous node,
// their lengths, and lpc.
struct Corner {
  D2 lpc, np, nm;
  double lp, lm;
};
__device__ __forceinline__ Corner cornerGeom(const Opts& o, long j, long i, int s) {
  const long jj[4] = {j, j, j + 1, j + 1}, ii[4] = {i, i + 1, i + 1, i};
  const D2 xp = eulerNode(o, jj[s], ii[s]);
  const D2 xpp = eulerNode(o, jj[(s + 1) & 3], ii[(s + 1) & 3]);
  const D2 xpm = eulerNode(o, jj[(s + 3) & 3], ii[(s + 3) & 3]);
  Corner g;
  g.np = {0.5 * (xpp.y - xp.y), 0.5 * (xp.x - xpp.x)};
  g.lp = d2norm(g.np);
  g.np = d2div(g.np, g.lp);
  g.nm = {0.5 * (xp.y - xpm.y), 0.5 * (xpm.x - xp.x)};
  g.lm = d2norm(g.nm);
  g.nm = d2div(g.nm, g.lm);
  g.lpc = d2add(d2mul(g.np, g.lp), d2mul(g.nm, g.lm));
  return g;
}
// the corner's dissipation geometry (computeDissipationMatrix,
// PhaseLagrange.cc:266-296)
__device__ __forceinline__ M22 cornerMatrix(const Corner& g) {
  return madd(mscale(g.lp, tens(g.np, g.np)), mscale(g.lm, tens(g.nm, g.nm)));
}

// start-of-step position of node (j, i): the Euler node, or the moving mesh
__device__ __forceinline__ D2 nodePos(const KP& kp, long j, long i) {
  if (kp.o.mode == MODE_LAGRANGE) {
    const long p = j * (kp.o.nx + 1) + i;
    return {kp.f.X[kp.cur][p], kp.f.Y[kp.cur][p]};
  }
  return eulerNode(kp.o, j, i);
}

// What the extrapolation of a corner of cell c needs from the cell.
template <int M>
struct CellState {
  double pk[M], fk[M], rk[M];
  D2 v, xc;
  M22 gv;
  D2 gp[M];
  double p, rho, c;
};
template <int M>
__device__ __forceinline__ CellState<M> loadCell(const KP& kp, const EucFields& e, long c) {
  const auto& f = kp.f;
  const int cur = kp.cur;
  CellState<M> s;
  const double V = e.ve[cur][c];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    s.pk[k] = f.p_k[0][k][c];
    s.fk[k] = f.f_k[k][c];
    s.rk[k] = f.rho_k[cur][k][c];
    // computeGradients (PhaseLagrange.cc:231-240): the material corner
    // forces of the last step over the cell volume
    s.gp[k] = d2div(D2{e.fsx[k][c], e.fsy[k][c]}, V);
  }
  s.v = {e.vx[cur][c], e.vy[cur][c]};
  s.xc = {e.xc[cur][c], e.yc[cur][c]};
  s.gv = {e.gv[0][c], e.gv[1][c], e.gv[2][c], e.gv[3][c]};
  s.p = f.p[0][c];
  s.rho = f.rho[cur][c];
  s.c = f.c[0][c];
  return s;
}

// range of the cells of a node (extrapolateValue, PhaseLagrange.cc:330-418):
// min from DBL_MAX, max from DBL_MIN (the smallest positive double, as
// numeric_limits<double>::min() in the reference)
template <int M>
struct NodeRange {
  double pmin[M], pmax[M];
  D2 vmin, vmax;
};

// Extrapolation to the corner at node position X of cell state s
// (extrapolateValue, PhaseLagrange.cc:322-478): the material pressures and
// the velocity, clamped to the node's range, and the mean pressure.
template <int M>
struct Extrap {
  double pe[M];
  double p;
  D2 v;
};
template <int M>
__device__ __forceinline__ Extrap<M> extrapolate(const Opts& o, const CellState<M>& s, D2 X, const NodeRange<M>& r) {
  Extrap<M> x;
  if (o.space_order != 2) {
#pragma unroll
    for (int k = 0; k < M; ++k) x.pe[k] = s.pk[k];
    x.p = s.p;
    x.v = s.v;
    return x;
  }
  const D2 d = d2sub(X, s.xc);
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const double ptmp = s.pk[k] + d2dot(s.gp[k], d);
    x.pe[k] = smax(smin(r.pmax[k], ptmp), r.pmin[k]);
  }
  x.p = 0.;
#pragma unroll
  for (int k = 0; k < M; ++k) x.p += s.fk[k] * x.pe[k];
  const D2 vt = d2add(s.v, matVec(s.gv, d));
  x.v = {smax(smin(r.vmax.x, vt.x), r.vmin.x), smax(smin(r.vmax.y, vt.y), r.vmin.y)};
  return x;
}

// nodeVelocityBoundaryCondition (ConditionsLimites.cc:109-123)
__device__ __forceinline__ D2 nodeBC(int bc, D2 b, M22 Mp, D2 G) {
  if (bc == MHY_BC_SYMMETRY) return d2mul(b, d2dot(G, b) / d2dot(matVec(Mp, b), b));
  if (bc == MHY_BC_IMPOSED_VELOCITY) return b;
  if (bc == MHY_BC_FREE_SURFACE) return matVec(inverse2(Mp), G);
  return {0.0, 0.0};
}
// nodeVelocityBoundaryConditionCorner (ConditionsLimites.cc:125-145)
__device__ __forceinline__ D2 nodeBCCorner(int bc1, D2 b1, int bc2, D2 b2, M22 Mp, D2 G) {
  if (bc1 == MHY_BC_SYMMETRY && bc2 == MHY_BC_SYMMETRY) {
    if (fabs(fabs(d2dot(b1, b2)) - d2norm(b1) * d2norm(b2)) < 1.0E-8)
      return d2mul(b1, d2dot(G, b1) / d2dot(matVec(Mp, b1), b1));
    return {0.0, 0.0};
  }
  if (bc1 == MHY_BC_IMPOSED_VELOCITY && bc2 == MHY_BC_IMPOSED_VELOCITY) return d2mul(d2add(b1, b2), 0.5);
  if (bc1 == MHY_BC_FREE_SURFACE && bc2 == MHY_BC_FREE_SURFACE) return matVec(inverse2(Mp), G);
  return {0.0, 0.0};
}

// EOS of the materials and mean pressure / sound speed of cell c
// (computeEOS PhaseLagrange.cc:70-105, computePressionMoyenne :139-154):
// the sound speed keeps its previous value in the max unless the density is
// positive, as in the reference (idempotent: the dumps run it once more).
template <int M>
__device__ __forceinline__ void eosCell(const KP& kp, long c) {
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const int cur = kp.cur;
  double p = 0.;
  double cs = f.c[0][c];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const double g = o.gammap[k];
    const double rk = f.rho_k[cur][k][c];
    double pk, ck;
    if (o.eos[k] == MHY_EOS_PERFECT_GAS) {
      pk = (g - 1.0) * rk * f.e_k[cur][k][c];
      ck = (rk > 0.) ? sqrt(g * pk / rk) : 1.e-20;
    } else {
      pk = 0.;
      ck = 1.e-20;
    }
    f.p_k[0][k][c] = pk;
    f.c_k[0][k][c] = ck;
    p += f.f_k[k][c] * pk;
    cs = smax(cs, ck);
  }
  const double rho = f.rho[cur][c];
  if (rho > 0.) cs = sqrt(o.gammap[0] * p / rho);
  f.p[0][c] = p;
  f.c[0][c] = cs;
}

__device__ __forceinline__ double waveMinD(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = smin(v, __shfl_xor(v, off, 64));
  return v;
}

}  // namespace

// ------------------------------------------------------------------ kernels

// State-of-step cell work: EOS, mean pressure, velocity gradient
// (computeGradients, PhaseLagrange.cc:199-230), cell masses (computeMass
// :251-260) and the cell time step (computem_cell_deltat :302-316) into the
// step's minimum (one 64-bit atomic min per block, exact and order-free).
template <int M>
__global__ __launch_bounds__(256) void k_euc_start(KP kp, EucFields e) {
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const int cur = kp.cur;
  const long nx = o.nx;
  double dtmin = DBL_MAX;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < kp.ncell; c += (long)gridDim.x * blockDim.x) {
    const long j = (long)((unsigned)c / (unsigned)nx), i = c - j * nx;
    eosCell<M>(kp, c);
    const double V = e.ve[cur][c];
    if (o.space_order == 2) {
      long jj[4], ii[4], p[4];
      quadNodes(nx, j, i, jj, ii, p);
      M22 r = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const Corner g = cornerGeom(o, j, i, s);
        r = madd(r, tens(D2{f.ux[cur][p[s]], f.uy[cur][p[s]]}, g.lpc));
      }
      e.gv[0][c] = r.a / V, e.gv[1][c] = r.b / V, e.gv[2][c] = r.c / V, e.gv[3][c] = r.d / V;
    }
    const double Mc = f.rho[cur][c] * V;
    f.M[c] = Mc;
#pragma unroll
    for (int k = 0; k < M; ++k) f.M_k[k][c] = f.mu_k[k][c] * Mc;
    const double cs = f.c[0][c];
    double cdt;
    if (o.mode == MODE_FULL) cdt = V / (e.perim[c] * (d2norm(D2{e.vx[cur][c], e.vy[cur][c]}) + cs));
    else cdt = V / (e.perim[c] * cs);
    dtmin = smin(dtmin, cdt);
  }
  __shared__ double smn[4];
  dtmin = waveMinD(dtmin);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) smn[w] = dtmin;
  __syncthreads();
  if (threadIdx.x == 0) {
    double lo = smn[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) lo = smin(lo, smn[k]);
    atomicMin(&f.sc->minKey[cur], dkeyPos(lo));
  }
}

// computedeltat (EucclhydRemap.cc:203-227) and updateTime: dt_{n+1} =
// min(cfl * minimum, 1.05 dt_n); below deltat_min the run stops with
// MHY_ERR_DT_MIN.  Resets the minimum of the next step's parity.
__global__ void k_euc_dt_finalize(KP kp) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Scalars* sc = kp.f.sc;
  const double lo = dkeyPosInv(sc->minKey[kp.cur]);
  const double dt1 = smin(kEucCfl * lo, sc->dt[kp.cur] * 1.05);
  sc->dt[kp.nxt] = dt1;
  sc->t[kp.nxt] = sc->t[kp.cur] + dt1;
  if (dt1 < kp.o.deltat_min) raiseErr(sc, MHY_ERR_DT_MIN);
  sc->minKey[kp.nxt] = dkeyPos(DBL_MAX);
}

// EOS and mean pressure only: the dumps at t_n print the pressure of state
// n, which the step forms at its start (called before the dump).
template <int M>
__global__ __launch_bounds__(256) void k_euc_eos(KP kp) {
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < kp.ncell; c += (long)gridDim.x * blockDim.x)
    eosCell<M>(kp, c);
}

// One thread per node: the node's ranges, G and the dissipation matrix from
// its cells in getCellsOfNode order (computeG PhaseLagrange.cc:484-504,
// computeNodeDissipationMatrixAndG :511-526), then the node velocity
// (computeBoundaryNodeVelocities ConditionsLimites.cc:19-107: each side's
// rule, the corners' rule at the four corners; computeNodeVelocity
// PhaseLagrange.cc:532-542 inside).  Without remap the node also moves
// (computeLagrangePosition, the new mesh of the next step).
template <int M>
__global__ __launch_bounds__(256) void k_euc_node(KP kp, EucFields e) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kp.nnode) return;
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const int cur = kp.cur;
  const long nx = o.nx, ny = o.ny;
  const long j = (long)((unsigned)p / (unsigned)(nx + 1)), i = p - j * (nx + 1);
  const bool hc[4] = {j < ny && i < nx, j < ny && i > 0, j > 0 && i < nx, j > 0 && i > 0};
  const long cj[4] = {j, j, j - 1, j - 1}, ci[4] = {i, i - 1, i, i - 1};
  const int slot[4] = {0, 1, 3, 2};  // the node's corner in each cell
  NodeRange<M> r;
  if (o.space_order == 2) {
    const double big = DBL_MAX, tiny = DBL_MIN;
#pragma unroll
    for (int k = 0; k < M; ++k) r.pmin[k] = big, r.pmax[k] = tiny;
    r.vmin = {big, big}, r.vmax = {tiny, tiny};
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (hc[q]) {
        const long c = cj[q] * nx + ci[q];
#pragma unroll
        for (int k = 0; k < M; ++k) r.pmin[k] = smin(r.pmin[k], (double)f.p_k[0][k][c]);
      }
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (hc[q]) {
        const long c = cj[q] * nx + ci[q];
#pragma unroll
        for (int k = 0; k < M; ++k) r.pmax[k] = smax(r.pmax[k], (double)f.p_k[0][k][c]);
      }
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (hc[q]) r.vmin.x = smin(r.vmin.x, e.vx[cur][cj[q] * nx + ci[q]]);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (hc[q]) r.vmax.x = smax(r.vmax.x, e.vx[cur][cj[q] * nx + ci[q]]);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (hc[q]) r.vmin.y = smin(r.vmin.y, e.vy[cur][cj[q] * nx + ci[q]]);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (hc[q]) r.vmax.y = smax(r.vmax.y, e.vy[cur][cj[q] * nx + ci[q]]);
#pragma unroll
    for (int k = 0; k < M; ++k) e.pmin[k][p] = r.pmin[k], e.pmax[k][p] = r.pmax[k];
    e.vmin[0][p] = r.vmin.x, e.vmin[1][p] = r.vmin.y, e.vmax[0][p] = r.vmax.x, e.vmax[1][p] = r.vmax.y;
  }
  const D2 X = nodePos(kp, j, i);
  D2 G = {0.0, 0.0};
  M22 Mp = {0.0, 0.0, 0.0, 0.0};
  for (int q = 0; q < 4; ++q) {
    if (!hc[q]) continue;
    const long c = cj[q] * nx + ci[q];
    const CellState<M> s = loadCell<M>(kp, e, c);
    const Corner g = cornerGeom(o, cj[q], ci[q], slot[q]);
    const Extrap<M> x = extrapolate<M>(o, s, X, r);
    const M22 Mpc = mscale(s.rho * s.c, cornerMatrix(g));
    G = d2add(G, d2add(matVec(Mpc, x.v), d2mul(g.lpc, x.p)));
    Mp = madd(Mp, Mpc);
  }
  const D2 bl = {o.bcv_left[0], o.bcv_left[1]}, br = {o.bcv_right[0], o.bcv_right[1]};
  const D2 bb = {o.bcv_bottom[0], o.bcv_bottom[1]}, bt = {o.bcv_top[0], o.bcv_top[1]};
  D2 v;
  const bool left = (i == 0), right = (i == nx), bottom = (j == 0), top = (j == ny);
  if (top && left) v = nodeBCCorner(o.bc_top, bt, o.bc_left, bl, Mp, G);
  else if (top && right) v = nodeBCCorner(o.bc_top, bt, o.bc_right, br, Mp, G);
  else if (bottom && left) v = nodeBCCorner(o.bc_bottom, bb, o.bc_left, bl, Mp, G);
  else if (bottom && right) v = nodeBCCorner(o.bc_bottom, bb, o.bc_right, br, Mp, G);
  else if (left) v = nodeBC(o.bc_left, bl, Mp, G);
  else if (right) v = nodeBC(o.bc_right, br, Mp, G);
  else if (top) v = nodeBC(o.bc_top, bt, Mp, G);
  else if (bottom) v = nodeBC(o.bc_bottom, bb, Mp, G);
  else v = matVec(inverse2(Mp), G);
  f.ux[kp.nxt][p] = v.x;
  f.uy[kp.nxt][p] = v.y;
  if (o.mode == MODE_LAGRANGE) {
    const double dt = f.sc->dt[cur];
    f.X[kp.nxt][p] = X.x + v.x * dt;
    f.Y[kp.nxt][p] = X.y + v.y * dt;
  }
}

// One thread per cell: the four corner forces (computeSubCellForce,
// PhaseLagrange.cc:625-666) from the corner extrapolations formed again with
// the node ranges, the Lagrange volume and centroid
// (computeLagrangeVolumeAndCenterOfGravity :671-715), then
// updateCellCenteredLagrangeVariables (:746-1011): Lagrange density,
// velocity and material energies, and the remap vector U_L (3m material
// components, the momentum, the kinetic energy of a conservative
// projection).  The sum of the material corner forces is kept for the next
// step's pressure gradients.  Without remap the cell writes state n+1.
template <int M>
__global__ __launch_bounds__(256) void k_euc_cell(KP kp, EucFields e) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= kp.ncell) return;
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const int cur = kp.cur, nxt = kp.nxt;
  const long nx = o.nx;
  const long j = (long)((unsigned)c / (unsigned)nx), i = c - j * nx;
  const double dt = f.sc->dt[cur], thr = o.threshold;
  long jj[4], ii[4], pn[4];
  quadNodes(nx, j, i, jj, ii, pn);
  const CellState<M> s = loadCell<M>(kp, e, c);
  D2 Xs[4], XL[4], vp[4], F[4], Fk[4][M];
  double r2 = 0.0;
  D2 r3 = {0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const long p = pn[q];
    NodeRange<M> r;
    if (o.space_order == 2) {
#pragma unroll
      for (int k = 0; k < M; ++k) r.pmin[k] = e.pmin[k][p], r.pmax[k] = e.pmax[k][p];
      r.vmin = {e.vmin[0][p], e.vmin[1][p]}, r.vmax = {e.vmax[0][p], e.vmax[1][p]};
    }
    Xs[q] = nodePos(kp, jj[q], ii[q]);
    vp[q] = {f.ux[nxt][p], f.uy[nxt][p]};
    XL[q] = d2add(Xs[q], d2mul(vp[q], dt));
    const Corner g = cornerGeom(o, j, i, q);
    const Extrap<M> x = extrapolate<M>(o, s, Xs[q], r);
    const M22 cm = cornerMatrix(g);
    const D2 dv = d2sub(vp[q], x.v);
    F[q] = d2add(d2mul(g.lpc, -x.p), matVec(mscale(s.rho * s.c, cm), dv));
#pragma unroll
    for (int k = 0; k < M; ++k) Fk[q][k] = d2add(d2mul(g.lpc, -x.pe[k]), matVec(mscale(s.rk[k] * s.c, cm), dv));
    r2 = r2 + d2dot(g.lpc, vp[q]);
    r3 = d2add(r3, F[q]);
  }
  // the next step's material pressure gradients (times the volume)
#pragma unroll
  for (int k = 0; k < M; ++k) {
    D2 sk = {0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 4; ++q) sk = d2add(sk, Fk[q][k]);
    e.fsx[k][c] = sk.x, e.fsy[k][c] = sk.y;
  }
  const double Mc = f.M[c];
  const double rhoL = 1 / (1 / s.rho + dt / Mc * r2);
  const D2 vL = d2add(s.v, d2div(d2mul(r3, dt), Mc));
  double pr4[M];
#pragma unroll
  for (int k = 0; k < M; ++k) pr4[k] = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const D2 w = d2sub(vp[q], d2mul(d2add(s.v, vL), 0.5));
#pragma unroll
    for (int k = 0; k < M; ++k) pr4[k] = pr4[k] + d2dot(Fk[q][k], w);
  }
  double peL[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const double Mk = f.M_k[k][c];
    peL[k] = 0.;
    if (s.fk[k] > thr && Mk != 0.) peL[k] = f.e_k[cur][k][c] + s.fk[k] * dt / Mk * pr4[k];
  }
  // Lagrange volume and centroid
  double r6 = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) r6 = r6 + d2cross(XL[q], XL[(q + 1) & 3]);
  const double vol = 0.5 * r6;
  D2 r7 = {0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const D2 a = XL[q], b = XL[(q + 1) & 3];
    r7 = d2add(r7, d2mul(d2add(a, b), d2cross(a, b)));
  }
  const D2 xcL = d2mul(r7, 1.0 / (6.0 * vol));
  if (o.mode == MODE_FULL) {
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const double mu = f.mu_k[k][c];
      f.UL[k][c] = s.fk[k] * vol;
      f.UL[M + k][c] = mu * vol * rhoL;
      f.UL[2 * M + k][c] = mu * vol * rhoL * peL[k];
    }
    f.UL[3 * M][c] = vol * rhoL * vL.x;
    f.UL[3 * M + 1][c] = vol * rhoL * vL.y;
    if (o.conservative) f.UL[3 * M + 2][c] = 0.5 * vol * rhoL * (vL.x * vL.x + vL.y * vL.y);
    f.VL[c] = vol;
    f.XcL[c] = xcL.x;
    f.YcL[c] = xcL.y;
  } else {
    // without remap: state n+1 (PhaseLagrange.cc:836-891) and the moved
    // mesh's volume and centroid; the perimeter of the start-of-step mesh
    // serves the next step's time step (computeLagrangePosition :601-616)
    e.vx[nxt][c] = vL.x, e.vy[nxt][c] = vL.y;
    double rho1 = 0., e1 = 0.;
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const double mu = f.mu_k[k][c], fk = s.fk[k];
      rho1 += mu * rhoL;
      f.rho_k[nxt][k][c] = (fk > thr) ? mu * rhoL / fk : 0.;
      e1 += mu * peL[k];
      f.e_k[nxt][k][c] = (fk > thr) ? peL[k] : 0.;
    }
    f.rho[nxt][c] = rho1;
    f.e[nxt][c] = e1;
    if (vL.x > 0.) e.outvx[c] = smax(vL.x, thr);
    if (vL.x < 0.) e.outvx[c] = smin(vL.x, -thr);
    if (vL.y > 0.) e.outvy[c] = smax(vL.y, thr);
    if (vL.y < 0.) e.outvy[c] = smin(vL.y, -thr);
    e.ve[nxt][c] = vol;
    e.xc[nxt][c] = xcL.x, e.yc[nxt][c] = xcL.y;
    double per = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) per = per + d2norm(d2sub(Xs[q], Xs[(q + 1) & 3]));
    e.perim[c] = per;
  }
}

// remapCellcenteredVariable (eucclhyd_remap/PhaseRemapFinal.cc:17-232) on
// one cell: fractions (small ones dropped, renormalised), mixed / pure
// flags, mass fractions, densities, velocity, specific energies (+ the
// kinetic-energy correction of a conservative projection), negative values
// zeroed, NaN raises MHY_ERR_NAN.
template <int M>
__global__ __launch_bounds__(256) void k_euc_final(KP kp, EucFields e) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= kp.ncell) return;
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const int nxt = kp.nxt;
  const double thr = o.threshold;
  double U[3 * M + 3];
#pragma unroll
  for (int q = 0; q < 3 * M + 2; ++q) U[q] = f.U2[q][c];
  U[3 * M + 2] = o.conservative ? f.U2[3 * M + 2][c] : 0.0;
  const double vol = e.ve[0][c];
  double volt = 0., masset = 0., vn[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    vn[k] = U[k];
    volt += vn[k];
    masset += U[M + k];
  }
  double vtn = 0.;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    vn[k] *= vol / volt;
    vtn += vn[k];
  }
  double fr[M], sf = 0.;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    fr[k] = vn[k] / vtn;
    if (fr[k] < thr) fr[k] = 0.;
    sf += fr[k];
  }
#pragma unroll
  for (int k = 0; k < M; ++k) fr[k] = fr[k] / sf;
  int matcell = 0, imatpure = -1;
#pragma unroll
  for (int k = 0; k < M; ++k)
    if (fr[k] > 0.) matcell++, imatpure = k;
  f.mixte[c] = (matcell > 1) ? 1 : 0;
  f.pure[c] = (matcell > 1) ? -1 : imatpure;
  double mf[M], fm = 0.;
#pragma unroll
  for (int k = 0; k < M; ++k) mf[k] = U[M + k] / masset;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    if (fr[k] < thr) mf[k] = 0.;
    fm += mf[k];
  }
#pragma unroll
  for (int k = 0; k < M; ++k) mf[k] /= fm;
  double dens[M], dn = 0.;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    dens[k] = 0.0;
    if (fr[k] > thr) dens[k] = U[M + k] / vn[k];
    dn += fr[k] * dens[k];
  }
  const D2 v = {U[3 * M] / (dn * vol), U[3 * M + 1] / (dn * vol)};
  double pe[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    pe[k] = 0.0;
    if ((fr[k] > thr) && (U[M + k] != 0.)) pe[k] = U[2 * M + k] / U[M + k];
  }
  double dec = 0.;
  if (o.conservative) dec = U[3 * M + 2] / masset - 0.5 * (v.x * v.x + v.y * v.y);
  double ek[M], en = 0.;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    ek[k] = pe[k];
    ek[k] += dec;
    en += mf[k] * ek[k];
  }
  double rk[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    rk[k] = dens[k];
    if (pe[k] < 0. || dens[k] < 0.) rk[k] = 0., ek[k] = 0., mf[k] = 0., fr[k] = 0.;
  }
  if (en != en || dn != dn) raiseErr(f.sc, MHY_ERR_NAN);
  f.rho[nxt][c] = dn;
  f.e[nxt][c] = en;
  e.vx[nxt][c] = v.x, e.vy[nxt][c] = v.y;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    f.f_k[k][c] = fr[k];
    f.mu_k[k][c] = mf[k];
    f.rho_k[nxt][k][c] = rk[k];
    f.e_k[nxt][k][c] = ek[k];
  }
  if (v.x > 0.) e.outvx[c] = smax(v.x, thr);
  if (v.x < 0.) e.outvx[c] = smin(v.x, -thr);
  if (v.y > 0.) e.outvy[c] = smax(v.y, thr);
  if (v.y < 0.) e.outvy[c] = smin(v.y, -thr);
}

#define MHY_EUC_INST(M)                                                  \
  template __global__ void k_euc_start<M>(KP, EucFields);               \
  template __global__ void k_euc_eos<M>(KP);                            \
  template __global__ void k_euc_node<M>(KP, EucFields);                \
  template __global__ void k_euc_cell<M>(KP, EucFields);                \
  template __global__ void k_euc_final<M>(KP, EucFields);
MHY_EUC_INST(1)
MHY_EUC_INST(2)
MHY_EUC_INST(3)
#undef MHY_EUC_INST

}  // namespace mhy
