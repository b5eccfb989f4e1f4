// Lagrange phase of the VNR scheme on CDNA4: CFL time-step / max-pressure
// reduction, nodal momentum update with boundary conditions, and the cell
// update (sub-volumes, densities, artificial viscosity, energy, perfect-gas
// EOS) fused with the preparation of the remap vectors.
//
// Arithmetic follows the reference operation by operation (compiled with
// -ffp-contract=off), so FP64 results are bitwise those of the CPU path.
#include "common.hpp"
#include "kernels.hpp"

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

__device__ __forceinline__ void raise(Scalars* sc, int code) { atomicCAS(&sc->err, 0, code); }

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
    const long j = c / nx, i = c - j * nx;
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
__device__ __forceinline__ void lagrangeNodeBody(const KPT<R>& kp, unsigned b) {
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const long p = (long)b * blockDim.x + threadIdx.x;
  if (p >= kp.nnode) return;
  const long nx = o.nx, ny = o.ny;
  const long j = p / (nx + 1), i = p - j * (nx + 1);
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
__global__ __launch_bounds__(256) void k_lagrange_node(KPT<R> kp) {
  lagrangeNodeBody(kp, blockIdx.x);
}

// Cell kernel: updateCellPos, computeSubVol, updateRho (+ periodic cells),
// computeTau, computeDivU, computeArtificialViscosity, updateEnergy,
// computeEOS, computePressionMoyenne (PhaseLagrange.cc:214-485,
// ConditionsLimites.cc:186-209), then the cell part of the remap
// preparation: computeVariablesForRemap and computeCellQuantitesForRemap
// (PrepareRemap.cc:65-176).  Arithmetic is FP64 in both forms; the mixed
// form takes the remap masses from the FP64 material masses M_k (equal to
// mu_k V rho in exact arithmetic), so that mass stays conserved to FP64
// round-off whatever the FP32 density storage (periodic cells, whose
// density is reset to its initial value, keep mu_k V rho).
template <class R, int M>
__device__ __forceinline__ void lagrangeCellBody(const KPT<R>& kp, unsigned b) {
  const Opts& o = kp.o;
  const auto& f = kp.f;
  const long c = (long)b * blockDim.x + threadIdx.x;
  if (c >= kp.ncell) return;
  constexpr int m = M;  // material count (template: the loops unroll)
  const long nx = o.nx;
  const long j = c / nx, i = c - j * nx;
  const int cur = kp.cur, nxt = kp.nxt;
  const Scalars* sc = f.sc;
  const double dt1 = sc->dt[nxt];
  // the cell's step-start words, loaded up front: their uses below are
  // conditional, and unconditional loads can all be in flight together
  double fk[NMAT], Mk0[NMAT], rk0[NMAT], qk0[NMAT], ek0[NMAT], pk0[NMAT], mu0[NMAT];
  for (int k = 0; k < m; ++k) {
    fk[k] = f.f_k[k][c];
    Mk0[k] = f.M_k[k][c];
    rk0[k] = (double)f.rho_k[cur][k][c];
    qk0[k] = (double)f.q_k[cur][k][c];
    ek0[k] = (double)f.e_k[cur][k][c];
    pk0[k] = (double)f.p_k[cur][k][c];
    mu0[k] = (double)f.mu_k[k][c];
  }
  const double rhoS = (double)f.rho[cur][c], divuS = (double)f.divu[cur][c], cS = (double)f.c[cur][c];
  const double vol0 = f.vol[cur][c], Mc0 = f.M[c];
  // end-of-Lagrange node positions, node order BL, BR, TR, TL
  D2 X[4];
  {
    const long jj[4] = {j, j, j + 1, j + 1}, ii[4] = {i, i + 1, i + 1, i};
    for (int s = 0; s < 4; ++s) {
      const long p = jj[s] * (nx + 1) + ii[s];
      if (o.mode == MODE_LAGRANGE) {
        X[s] = {f.X[nxt][p], f.Y[nxt][p]};
      } else {
        const D2 e = eulerNode(o, jj[s], ii[s]);
        X[s] = {e.x + (double)f.ux[nxt][p] * dt1, e.y + (double)f.uy[nxt][p] * dt1};
      }
    }
  }
  D2 xc = {0.0, 0.0};
  for (int s = 0; s < 4; ++s) xc = d2add(xc, X[s]);
  xc = d2mul(xc, 0.25);
  double V = 0.0;
  for (int s = 0; s < 4; ++s) V = V + subVolume(xc, X[(s + 3) & 3], X[s], X[(s + 1) & 3]);

  double rk1[NMAT];
  double rho1 = 0.0;
  for (int k = 0; k < m; ++k) {
    // absent materials keep the stale buffer value in the reference; their
    // volume fraction is exactly zero, so the value never contributes
    rk1[k] = (fk[k] > o.threshold) ? Mk0[k] / (fk[k] * V)
                                   : (o.mode == MODE_FULL ? 0.0 : (double)f.rho_k[nxt][k][c]);
    rho1 += fk[k] * rk1[k];
  }
  const bool periodicCell = o.periodic_right && (i == 0 || i == nx - 1);
  if (periodicCell) {
    const long L = j * nx;  // updatePeriodicBoundaryConditions
    rho1 = (double)f.rho0[L];
    for (int k = 0; k < m; ++k) rk1[k] = f.rho_k0[k][L];
  }

  double q1 = 0.0, e1 = 0.0;
  if (o.mode != MODE_ADVECTION) {
    // mixed form: the step-start densities from the FP64 masses and volume,
    // so that 1/rho1 - 1/rho0 (divided by dt in div u) carries no FP32
    // rounding of the stored density
    const double rho0 = kMixed<R> ? Mc0 / vol0 : rhoS;
    const double tau1 = 0.5 * (1.0 / rho1 + 1.0 / rho0);
    const double div1 = 1.0 / dt1 * (1.0 / rho1 - 1.0 / rho0) / tau1;
    if (divuS < 0.0) {
      const double r0 = vol0;
      q1 = 1.0 / tau1 * (-0.5 * sqrt(r0) * cS * div1 + (o.gamma + 1) / 2.0 * r0 * div1 * div1);
    } else {
      q1 = 0.0;
    }
    const double P = sc->pmax;
    if (q1 > P) q1 = P;
    double qk1[NMAT], ek1[NMAT], pk1[NMAT], ck1[NMAT];
    for (int k = 0; k < m; ++k) qk1[k] = fk[k] * q1;
    bool bad = false;
    for (int k = 0; k < m; ++k) {
      ek1[k] = 0.0;
      const double r1 = rk1[k];
      const double r0 = (kMixed<R> && fk[k] > o.threshold) ? Mk0[k] / (fk[k] * vol0) : rk0[k];
      if ((r1 > o.threshold) && (r0 > o.threshold)) {
        const double qn = qk0[k];
        double pseudo = 0.0;
        if ((o.pseudo_centree == 1) && ((qk1[k] + qn) * (1.0 / r1 - 1.0 / r0) > 0.)) pseudo = 0.5 * (qk1[k] + qn);
        if (o.pseudo_centree == 0) pseudo = qk1[k];
        const double den = 1 + 0.5 * (o.gammap[k] - 1.0) * r1 * (1.0 / r1 - 1.0 / r0);
        const double num = ek0[k] - (0.5 * pk0[k] + pseudo) * (1.0 / r1 - 1.0 / r0);
        ek1[k] = num / den;
        e1 += mu0[k] * ek1[k];
        if (fabs(ek1[k]) > 5. && o.abort_energy) bad = true;
      }
    }
    if (bad) raise(f.sc, MHY_ERR_ENERGY);
    // computeEOS: perfect gas and void (PhaseLagrange.cc:392-430)
    for (int k = 0; k < m; ++k) {
      const double g = o.gammap[k];
      if (o.eos[k] == MHY_EOS_PERFECT_GAS) {
        pk1[k] = (g - 1.0) * rk1[k] * ek1[k];
        ck1[k] = sqrt(g * (g - 1.0) * ek1[k]);
      } else if (o.eos[k] == MHY_EOS_VOID) {
        pk1[k] = pk0[k];
        ck1[k] = f.c_k[cur][k][c];
        ek1[k] = ek0[k];
        e1 = 0.0;
      } else {
        pk1[k] = f.p_k[nxt][k][c];
        ck1[k] = f.c_k[nxt][k][c];
      }
    }
    // computePressionMoyenne (PhaseLagrange.cc:464-485)
    double p1 = 0.0, c1 = 0.0;
    const bool needStale = !(rho1 > 0.);
    if (needStale) c1 = (double)f.c[nxt][c];
    for (int k = 0; k < m; ++k) {
      p1 += fk[k] * pk1[k];
      if (needStale) c1 = smax(c1, ck1[k]);
    }
    if (rho1 > 0.) c1 = sqrt(o.gammap[0] * (o.gammap[0] - 1.0) * e1);
    for (int k = 0; k < m; ++k)
      if (o.eos[k] == MHY_EOS_VOID) e1 += mu0[k] * ek1[k];
    f.divu[nxt][c] = div1;
    for (int k = 0; k < m; ++k) f.q_k[nxt][k][c] = qk1[k];
    if (o.mode == MODE_LAGRANGE) {
      f.rho[nxt][c] = rho1;
      f.q[nxt][c] = q1;
      f.e[nxt][c] = e1;
      f.p[nxt][c] = p1;
      f.c[nxt][c] = c1;
      for (int k = 0; k < m; ++k) {
        f.rho_k[nxt][k][c] = rk1[k];
        f.e_k[nxt][k][c] = ek1[k];
        f.p_k[nxt][k][c] = pk1[k];
        f.c_k[nxt][k][c] = ck1[k];
      }
      f.vol[nxt][c] = V;  // node_cellvolume_n swap (VnrRemap.cc:217-221)
      return;
    }
  } else {
    // advection: the energy swap of VnrRemap.cc:170 makes the remap read the
    // current energy; the pseudo buffer of the n+1 slot is read as left by
    // the previous remap of that slot
    e1 = (double)f.e[cur][c];
    q1 = (double)f.q[nxt][c];
  }
  // computeVariablesForRemap (PrepareRemap.cc:96-176), live components only
  for (int k = 0; k < m; ++k) {
    f.UL[k][c] = fk[k] * V;
    if (kMixed<R> && !periodicCell) {
      const double Mk = Mk0[k];
      f.UL[m + k][c] = Mk;
      f.UL[2 * m + k][c] = Mk * e1;
    } else {
      const double mu = mu0[k];
      f.UL[m + k][c] = mu * V * rho1;
      f.UL[2 * m + k][c] = mu * V * rho1 * e1;
    }
  }
  f.UL[3 * m][c] = q1 * V;
  f.VL[c] = V;
  if (o.methode == 1) f.rL[c] = rho1;
  if (o.pp == 1) f.qL[c] = q1;
  // computeCellQuantitesForRemap (PrepareRemap.cc:65-90): polygon centroid
  D2 r = {0.0, 0.0};
  for (int s = 0; s < 4; ++s) {
    const D2 a = X[s], b = X[(s + 1) & 3];
    r = d2add(r, d2mul(d2add(a, b), d2cross(a, b)));
  }
  r = d2mul(r, 1.0 / (6.0 * V));
  f.XcL[c] = r.x;
  f.YcL[c] = r.y;
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
