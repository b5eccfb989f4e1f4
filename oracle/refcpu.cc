/*
 * ORACLE -- test infrastructure, not product code.
 *
 * CPU restatement of MaHyCo's VNR Lagrange + directional remap time loop,
 * and of the Eucclhyd Lagrange scheme that calls the same cell sweeps: the
 * same arithmetic, in the same operation order, on the same n/n+1 buffer
 * swaps as the reference, but with closed-form structured indexing instead
 * of the reference's O(N) mesh queries.  Each function cites the reference
 * file:line it restates.  Built with g++ -ffp-contract=off so no
 * multiply-add is fused (the reference is compiled for x86-64 without FMA).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, as the checker; the product never calls it.
 *
 * Pinned by: the 63 golden VTP fixtures of the reference (37 VNR and 26
 * Eucclhyd tests/<case>/reference/ VTP files, compared as text), plus
 * closed forms and invariants (limiters, swept integrals, PenteBorne-mixte
 * profiles, mass conservation of the sweeps, uniform-state preservation,
 * X/Y mirror symmetry) in tests/.
 */
#include "refcpu.h"

#include <sys/stat.h>
#include <sys/types.h>

#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <map>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int NMAT = MHY_NBMATMAX;
constexpr int NEQ = 3 * NMAT + 2 + 1 + 1;  // includes/Constantes.h:13

typedef std::array<double, 2> V2;
typedef std::array<double, NMAT> VM;
typedef std::array<double, NEQ> VE;
typedef std::array<double, 4> V4;

// std::min / std::max argument conventions (NaN and signed-zero behaviour
// depend on the argument order): MathFunctions.h:23-24.
inline double smin(double a, double b) { return (b < a) ? b : a; }
inline double smax(double a, double b) { return (a < b) ? b : a; }

// includes/Freefunctions.h:86-92
inline double divideNoExcept(double a, double b) {
  if (std::fabs(b) < 1.0E-12) return 0.0;
  return a / b;
}
// includes/Freefunctions.h:13-23 (accumulates from 0.0, left to right)
inline double dot2(const V2& a, const V2& b) {
  double r = 0.0;
  r = r + a[0] * b[0];
  r = r + a[1] * b[1];
  return r;
}
inline double norm2(const V2& a) { return std::sqrt(dot2(a, a)); }
inline double cross2(const V2& a, const V2& b) { return a[0] * b[1] - a[1] * b[0]; }
inline V2 vadd(const V2& a, const V2& b) { return {a[0] + b[0], a[1] + b[1]}; }
inline V2 vsub(const V2& a, const V2& b) { return {a[0] - b[0], a[1] - b[1]}; }
inline V2 vmul(const V2& a, double s) { return {a[0] * s, a[1] * s}; }
inline V2 vdiv(const V2& a, double s) { return {a[0] / s, a[1] / s}; }
inline VE ezero() {
  VE z;
  z.fill(0.0);
  return z;
}

// includes/Freefunctions.h:25-58: |C| . (C/|C|) per half edge.
inline double lpcPlus(const V2& xp, const V2& xpPlus) {
  V2 n = {0.5 * (xpPlus[1] - xp[1]), 0.5 * (xp[0] - xpPlus[0])};
  return norm2(n);
}
inline V2 npcPlus(const V2& xp, const V2& xpPlus) {
  V2 n = {0.5 * (xpPlus[1] - xp[1]), 0.5 * (xp[0] - xpPlus[0])};
  double l = norm2(n);
  return vdiv(n, l);
}
inline V2 lpcNpc(const V2& xp, const V2& xpPlus, const V2& xpMinus) {
  V2 a = vmul(npcPlus(xp, xpPlus), lpcPlus(xp, xpPlus));
  V2 b = vmul(npcPlus(xpMinus, xp), lpcPlus(xpMinus, xp));
  return vadd(a, b);
}
// includes/Freefunctions.h:60-65
inline V2 symmetricVector(const V2& v, const V2& sigma) {
  double d = 2.0 * dot2(v, sigma);
  return vsub(vmul(sigma, d), v);
}

/* Structured 2-D mesh with the numbering of
 * cppnabla/mesh/CartesianMesh2DGenerator.cc:26-101 and the queries of
 * cppnabla/mesh/CartesianMesh2D.cc, in closed form. */
struct Mesh {
  long nx = 0, ny = 0;
  long nbCells = 0, nbNodes = 0, nbFaces = 0;
  long node(long j, long i) const { return j * (nx + 1) + i; }
  long cell(long j, long i) const { return j * nx + i; }
  // horizontal edge on node row j starting at node column i
  long hface(long j, long i) const { return j < ny ? j * (2 * nx + 1) + 2 * i : ny * (2 * nx + 1) + i; }
  // vertical edge on node column i starting at node row j
  long vface(long j, long i) const { return i < nx ? j * (2 * nx + 1) + 2 * i + 1 : j * (2 * nx + 1) + 2 * nx; }
  void faceInfo(long f, bool& vert, long& j, long& i) const {
    long row = f / (2 * nx + 1), k = f - row * (2 * nx + 1);
    if (row < ny) {
      if (k == 2 * nx) {
        vert = true, j = row, i = nx;
      } else if (k % 2 == 0) {
        vert = false, j = row, i = k / 2;
      } else {
        vert = true, j = row, i = (k - 1) / 2;
      }
    } else {
      vert = false, j = ny, i = k;
    }
  }
  long firstNode(long f) const {
    bool v; long j, i;
    faceInfo(f, v, j, i);
    return node(j, i);
  }
  long secondNode(long f) const {
    bool v; long j, i;
    faceInfo(f, v, j, i);
    return v ? node(j + 1, i) : node(j, i + 1);
  }
  // getCellsOfFace order (CartesianMesh2D.cc:110-136): back = [0], front = [1]
  void cellsOfFace(long f, long& back, long& front) const {
    bool v; long j, i;
    faceInfo(f, v, j, i);
    front = -1;
    if (v) {
      if (i == 0) back = cell(j, 0);
      else if (i == nx) back = cell(j, nx - 1);
      else back = cell(j, i - 1), front = cell(j, i);
    } else {
      if (j == 0) back = cell(0, i);
      else if (j == ny) back = cell(ny - 1, i);
      else back = cell(j, i), front = cell(j - 1, i);
    }
  }
  long backCell(long f) const { long b, fr; cellsOfFace(f, b, fr); return b; }
  long frontCell(long f) const { long b, fr; cellsOfFace(f, b, fr); return fr; }
  bool innerNode(long p) const {
    long j = p / (nx + 1), i = p - j * (nx + 1);
    return i != 0 && j != 0 && i != nx && j != ny;
  }
  // CartesianMesh2D.cc:251-255: an edge is inner when one of its nodes is
  bool innerFace(long f) const { return innerNode(firstNode(f)) || innerNode(secondNode(f)); }
  bool isVertical(long f) const { bool v; long j, i; faceInfo(f, v, j, i); return v; }
  // getFacesOfCell slots {bottom, left, right, top} (CartesianMesh2D.cc:168-178)
  std::array<long, 4> facesOfCell(long c) const {
    long j = c / nx, i = c - j * nx;
    return {hface(j, i), vface(j, i), vface(j, i + 1), hface(j + 1, i)};
  }
  // quad node order BL, BR, TR, TL (Generator.cc:86-93)
  std::array<long, 4> nodesOfCell(long c) const {
    long j = c / nx, i = c - j * nx;
    return {node(j, i), node(j, i + 1), node(j + 1, i + 1), node(j + 1, i)};
  }
  // getCellsOfNode order (CartesianMesh2D.cc:97-108), with the corner slot
  // of the node inside each cell.
  int cellsOfNode(long p, long* cells, int* slot) const {
    long j = p / (nx + 1), i = p - j * (nx + 1);
    int k = 0;
    if (j < ny && i < nx) cells[k] = cell(j, i), slot[k++] = 0;
    if (j < ny && i > 0) cells[k] = cell(j, i - 1), slot[k++] = 1;
    if (j > 0 && i < nx) cells[k] = cell(j - 1, i), slot[k++] = 3;
    if (j > 0 && i > 0) cells[k] = cell(j - 1, i - 1), slot[k++] = 2;
    return k;
  }
  long leftNode(long p) const { long j = p / (nx + 1), i = p - j * (nx + 1); return i != 0 ? node(j, i - 1) : -1; }
  long rightNode(long p) const { long j = p / (nx + 1), i = p - j * (nx + 1); return i != nx ? node(j, i + 1) : -1; }
  long bottomNode(long p) const { long j = p / (nx + 1), i = p - j * (nx + 1); return j != 0 ? node(j - 1, i) : -1; }
  long topNode(long p) const { long j = p / (nx + 1), i = p - j * (nx + 1); return j != ny ? node(j + 1, i) : -1; }
  // remap/UtilesMesh.cc:5-51 (neighbour, or the cell itself at the edge)
  long leftCells(long c) const { long b = backCell(facesOfCell(c)[1]); return b == -1 ? c : b; }
  long rightCells(long c) const { long f = frontCell(facesOfCell(c)[2]); return f == -1 ? c : f; }
  long bottomCells(long c) const { long f = frontCell(facesOfCell(c)[0]); return f == -1 ? c : f; }
  long topCells(long c) const { long b = backCell(facesOfCell(c)[3]); return b == -1 ? c : b; }
};

// splitmix64: the initialisers' random stream
struct Rng {
  unsigned long long s;
  explicit Rng(unsigned long long seed) : s(seed) {}
  unsigned long long next() {
    unsigned long long z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

struct ref_ctx {
  mhy_config o;
  Mesh mesh;
  int status = MHY_OK;

  // boundary conditions: includes/ConditionsLimites.h:9-43
  int leftBC = 0, rightBC = 0, topBC = 0, bottomBC = 0;
  V2 leftBCValue = {0.0, 1.0}, rightBCValue = {0.0, 1.0}, topBCValue = {1.0, 0.0}, bottomBCValue = {1.0, 0.0};
  int leftCellBC = 0, rightCellBC = 0, bottomCellBC = 0, topCellBC = 0;
  int leftFluxBC = 0, rightFluxBC = 0, bottomFluxBC = 0, topFluxBC = 0, FluxBC = 0;
  VE leftFluxBCValue = ezero(), rightFluxBCValue = ezero(), bottomFluxBCValue = ezero(), topFluxBCValue = ezero();

  // time management: includes/GestionTemps.h:8-19
  double t_n = 0.0, t_nplus1 = 0.0, deltat_n = 0.0, deltat_nplus1 = 0.0;
  int n = 0, nbCalls = 0;
  double lastDump = 0.0;
  bool finished = false;

  // Vnr fields (vnr_remap/Vnr.h:101-174)
  std::vector<V2> node_coord_n, node_coord_nplus1, node_coord_n0;
  std::vector<V4> node_cellvolume_n, node_cellvolume_nplus1, node_cellvolume_n0;
  std::vector<double> euler_volume;
  std::vector<double> density_n, density_nplus1, density_n0;
  std::vector<VM> density_env_n, density_env_nplus1, density_env_n0;
  std::vector<double> pressure_n, pressure_nplus1, pressure_n0;
  std::vector<VM> pressure_env_n, pressure_env_nplus1, pressure_env_n0;
  std::vector<double> pseudo_n, pseudo_nplus1, pseudo_n0;
  std::vector<VM> pseudo_env_n, pseudo_env_nplus1, pseudo_env_n0;
  std::vector<double> tau_n, tau_nplus1, tau_n0;
  std::vector<VM> tau_env_n, tau_env_nplus1, tau_env_n0;
  std::vector<double> divu_n, divu_nplus1, divu_n0;
  std::vector<double> speed_n, speed_nplus1, speed_n0;
  std::vector<VM> speed_env_n, speed_env_nplus1, speed_env_n0;
  std::vector<double> energy_n, energy_nplus1, energy_n0;
  std::vector<VM> energy_env_n, energy_env_nplus1, energy_env_n0;
  std::vector<V2> node_velocity_n, node_velocity_nplus1, node_velocity_n0;
  std::vector<double> x_velocity, y_velocity, node_mass, cell_mass;
  std::vector<VM> cell_mass_env, mass_fraction_env, fracvol_env;
  std::vector<double> fracvol_env1, fracvol_env2, fracvol_env3;
  std::vector<V2> cell_coord_n, cell_coord_nplus1, cell_coord_n0;
  std::vector<std::array<V2, 4>> cqs;

  // VariablesLagRemap (includes/VariablesLagRemap.h:30-50)
  std::vector<VE> Phi, DualPhi, ULagrange, UDualLagrange, Uremap2, UDualremap2;
  std::vector<double> deltaxLagrange, faceLength, faceLengthLagrange, faceNormalVelocity, vLagrange, rLagrange;
  std::vector<V2> XLagrange, XcLagrange, Xf, faceNormal;
  std::vector<std::array<V2, 4>> outerFaceNormal;
  std::vector<int> mixte, pure;
  bool x_then_y_n = true, x_then_y_nplus1 = true;

  // Remap private fields (remap/Remap.h:66-99)
  std::vector<double> HvLagrange;
  std::vector<VE> Uremap1, UDualremap1, gradPhiFace1, gradPhiFace2, gradPhi1, gradPhi2, phiFace1, phiFace2;
  std::vector<VE> DualphiFlux1, DualphiFlux2, deltaPhiFaceAv, deltaPhiFaceAr, gradDualPhi1, gradDualPhi2;
  std::vector<std::array<VE, 4>> FluxFace1, FluxFace2;
  std::vector<double> RightFluxMasse, LeftFluxMasse, TopFluxMasse, BottomFluxMasse;
  std::vector<std::array<long, 2>> VerticalFaceOfNode, HorizontalFaceOfNode;
  std::vector<V2> TopupwindVelocity, BottomupwindVelocity, RightupwindVelocity, LeftupwindVelocity;

  // output
  std::string outdir;
  bool writerDisabled = true;
  std::map<double, std::string> fileNameByTimes;

  // ----------------------------------------------------------------------
  void allocate();
  void initBoundaryConditions();
  void initCellPos();
  int init();
  void initSubVol();
  void initMeshGeometryForFaces();
  void facesOfNode();
  void computeCellMass();
  void initDeltaT();
  void initInternalEnergy();
  void initPseudo();
  void setUpTimeLoopN();
  void computeNodeMass();
  int prologue();

  void computeCornerNormal();
  void computeDeltaT();
  void updateVelocity();
  void updateVelocityBoundaryConditions();
  void updatePosition();
  void updateCellPos();
  void computeSubVol();
  void updateRho();
  void updatePeriodicBoundaryConditions();
  void computeTau();
  void computeDivU();
  void computeArtificialViscosity();
  int updateEnergy();
  void computeEOS();
  void computePressionMoyenne();
  void computeVariablesForRemap();
  void computeCellQuantitesForRemap();
  void computeFaceQuantitesForRemap();
  void computeGradPhiFace(int sweep);
  void computeGradPhi(int sweep);
  void computeUpwindFaceQuantitiesForProjection(int sweep);
  void computeUremap(int sweep);
  void computeDualUremap(int sweep);
  int remapVariables();
  int iterate(bool dump, bool* continueLoop);
  void dumpVariables();
  void writeVtp(const std::string& module, long N, const V2* nodes, long C,
                const std::map<std::string, const double*>& cellVars,
                const std::map<std::string, const double*>& nodeVars, std::map<double, std::string>& byTimes);
  void swapAll();

  // Eucclhyd fields (eucclhyd_remap/Eucclhyd.h:86-160), allocated for that
  // scheme only.  The reference keeps the corner quantities per node and
  // cell of the node; here they are kept per cell corner, slot s in the
  // quad node order BL, BR, TR, TL (one-to-one, same values).
  typedef std::array<double, 4> M22;  // row-major 2x2
  std::vector<std::array<V2, 4>> e_lpc, e_nplus, e_nminus;
  std::vector<V4> e_lplus, e_lminus;
  std::vector<double> e_perimeter, e_cell_deltat, x_cell_velocity, y_cell_velocity;
  std::vector<V2> cell_velocity_n, cell_velocity_nplus1, cell_velocity_n0;
  std::vector<M22> e_velgrad;
  std::vector<std::array<V2, NMAT>> e_pgrad_env;
  std::vector<V4> e_p_extrap;
  std::vector<std::array<VM, 4>> e_pe_extrap;
  std::vector<std::array<V2, 4>> e_v_extrap, e_F_n, e_F_nplus1;
  std::vector<std::array<std::array<V2, NMAT>, 4>> e_Fenv_n, e_Fenv_nplus1;
  std::vector<std::array<M22, 4>> e_M;
  std::vector<std::array<std::array<M22, NMAT>, 4>> e_Menv;
  std::vector<M22> e_nodeM;
  std::vector<V2> e_nodeG, XfLagrange;
  std::vector<VM> lagrange_volume;
  std::map<double, std::string> particleFileNameByTimes;

  bool euc() const { return o.schema == MHY_SCHEMA_EUCCLHYD; }
  void eucAllocate();
  int eucPrologue();
  void eucInitBoundaryConditions();
  void eucCornerNormals();
  void eucInitMeshGeometryForCells();
  int eucInitState();
  void eucSetUpTimeLoopN();
  void eucComputeEOS();
  void eucPressionMoyenne();
  void eucGradients();
  void eucMass();
  void eucDissipationMatrix();
  void eucCellDeltat();
  void eucExtrapolate();
  void eucNodeG();
  void eucNodeDissipation();
  int eucDeltat();
  V2 eucNodeBC(int bc, const V2& bcv, const M22& Mp, const V2& Gp) const;
  V2 eucNodeBCCorner(int bc1, const V2& bcv1, int bc2, const V2& bcv2, const M22& Mp, const V2& Gp) const;
  void eucNodeVelocities();
  void eucFaceVelocity();
  void eucLagrangePosition();
  void eucSubCellForce();
  void eucLagrangeGeometry();
  void eucFaceDeltax();
  void eucCellUpdate();
  int eucRemapFinal();
  int eucIterate(bool dump, bool* continueLoop);
  void eucDump();
  void eucSwap();

  // remap helpers (remap/UtilesRemap.cc, remap/UtilesRemap-Impl.h)
  double fluxLimiter(int lim, double r) const;
  double fluxLimiterPP(int lim, double gp, double gm, double y0, double yp, double ym, double h0, double hp,
                       double hm) const;
  double computeY0(int lim, double y0, double yp, double ym, double h0, double hp, double hm, int type) const;
  double computexgxd(double y0, double yp, double ym, double h0, double y0p, double y0m, int type) const;
  double computeygyd(double y0, double yp, double ym, double h0, double y0p, double y0m, double g, int type) const;
  static double INT2Y(double X, double x0, double y0, double x1, double y1);
  VE computeAndLimitGradPhi(int lim, const VE& gp, const VE& gm, const VE& phi, const VE& phip, const VE& phim,
                            double h0, double hp, double hm) const;
  void computeFluxPP(bool pureVariant, const VE& grad, const VE& phi, const VE& phip, const VE& phim, double h0,
                     double hp, double hm, double fnv, double dt, int type, double thr, int complet, double dualv,
                     int calcDual, VE* pFlux, VE* pFluxDual) const;
  void gradPhiCell(const VE& gp, const VE& gm, const VE& phi, const VE& phip, const VE& phim, const int* mix3,
                   const int* pure3, double h0, double hp, double hm, double Fm, double Fp, double fdual, double dt,
                   VE* grad, VE* Ar, VE* Av, VE* dual) const;
  VE computeRemapFlux(double fnv, const V2& fn, double fl, const VE& phiFace, const V2& ofn, const V2& exy,
                      double dt) const;
  VE computeBoundaryFluxes(int proj, long c, const V2& exy) const;
  double ComputeFluxOrdre3(double ymmm, double ymm, double ym, double yp, double ypp, double yppp, double hmmm,
                           double hmm, double hm, double hp, double hpp, double hppp, double vdt) const;
  VE dualGradPhi(bool horizontal, long p) const;
  void massFluxes(int sweep, bool horizontal, long p);
  V2 upwindVelocity(int side, long nb, long p, const VE& gradNb, const VE& grad0) const;
};

namespace {
template <class T>
void zfill(std::vector<T>& v, size_t n) {
  T z;
  std::memset(&z, 0, sizeof(T));
  v.assign(n, z);
}
}  // namespace

void ref_ctx::allocate() {
  const size_t C = mesh.nbCells, N = mesh.nbNodes, F = mesh.nbFaces;
  zfill(node_coord_n, N), zfill(node_coord_nplus1, N), zfill(node_coord_n0, N);
  zfill(node_cellvolume_n, C), zfill(node_cellvolume_nplus1, C), zfill(node_cellvolume_n0, C);
  zfill(euler_volume, C);
  zfill(density_n, C), zfill(density_nplus1, C), zfill(density_n0, C);
  zfill(density_env_n, C), zfill(density_env_nplus1, C), zfill(density_env_n0, C);
  zfill(pressure_n, C), zfill(pressure_nplus1, C), zfill(pressure_n0, C);
  zfill(pressure_env_n, C), zfill(pressure_env_nplus1, C), zfill(pressure_env_n0, C);
  zfill(pseudo_n, C), zfill(pseudo_nplus1, C), zfill(pseudo_n0, C);
  zfill(pseudo_env_n, C), zfill(pseudo_env_nplus1, C), zfill(pseudo_env_n0, C);
  zfill(tau_n, C), zfill(tau_nplus1, C), zfill(tau_n0, C);
  zfill(tau_env_n, C), zfill(tau_env_nplus1, C), zfill(tau_env_n0, C);
  zfill(divu_n, C), zfill(divu_nplus1, C), zfill(divu_n0, C);
  zfill(speed_n, C), zfill(speed_nplus1, C), zfill(speed_n0, C);
  zfill(speed_env_n, C), zfill(speed_env_nplus1, C), zfill(speed_env_n0, C);
  zfill(energy_n, C), zfill(energy_nplus1, C), zfill(energy_n0, C);
  zfill(energy_env_n, C), zfill(energy_env_nplus1, C), zfill(energy_env_n0, C);
  zfill(node_velocity_n, N), zfill(node_velocity_nplus1, N), zfill(node_velocity_n0, N);
  zfill(x_velocity, N), zfill(y_velocity, N), zfill(node_mass, N), zfill(cell_mass, C);
  zfill(cell_mass_env, C), zfill(mass_fraction_env, C), zfill(fracvol_env, C);
  zfill(fracvol_env1, C), zfill(fracvol_env2, C), zfill(fracvol_env3, C);
  zfill(cell_coord_n, C), zfill(cell_coord_nplus1, C), zfill(cell_coord_n0, C);
  zfill(cqs, C);
  zfill(Phi, C), zfill(DualPhi, N), zfill(ULagrange, C), zfill(UDualLagrange, N), zfill(Uremap2, C);
  zfill(UDualremap2, N);
  zfill(deltaxLagrange, F), zfill(faceLength, F), zfill(faceLengthLagrange, F), zfill(faceNormalVelocity, F);
  zfill(vLagrange, C), zfill(rLagrange, C);
  zfill(XLagrange, N), zfill(XcLagrange, C), zfill(Xf, F), zfill(faceNormal, F);
  zfill(outerFaceNormal, C);
  zfill(mixte, C), zfill(pure, C);
  zfill(HvLagrange, C);
  zfill(Uremap1, C), zfill(UDualremap1, N), zfill(gradPhiFace1, F), zfill(gradPhiFace2, F);
  zfill(gradPhi1, C), zfill(gradPhi2, C), zfill(phiFace1, F), zfill(phiFace2, F);
  zfill(DualphiFlux1, C), zfill(DualphiFlux2, C), zfill(deltaPhiFaceAv, C), zfill(deltaPhiFaceAr, C);
  zfill(gradDualPhi1, N), zfill(gradDualPhi2, N);
  zfill(FluxFace1, C), zfill(FluxFace2, C);
  zfill(RightFluxMasse, N), zfill(LeftFluxMasse, N), zfill(TopFluxMasse, N), zfill(BottomFluxMasse, N);
  zfill(VerticalFaceOfNode, N), zfill(HorizontalFaceOfNode, N);
  zfill(TopupwindVelocity, N), zfill(BottomupwindVelocity, N), zfill(RightupwindVelocity, N);
  zfill(LeftupwindVelocity, N);
  if (euc()) eucAllocate();
  // node coordinates: CartesianMesh2DGenerator.cc:52-55 (x = dx*i, y = dy*j)
  for (long j = 0; j <= mesh.ny; ++j)
    for (long i = 0; i <= mesh.nx; ++i) {
      long p = mesh.node(j, i);
      node_coord_n0[p] = {o.dx * (double)(size_t)i, o.dy * (double)(size_t)j};
    }
}

// vnr_remap/Init.cc:9-95, extended with all-symmetry boundaries for the
// initialisers that are not in the reference.
void ref_ctx::initBoundaryConditions() {
  const int sym = MHY_BC_SYMMETRY, imp = MHY_BC_IMPOSED_VELOCITY;
  const V2 ex = {1.0, 0.0}, ey = {0.0, 1.0};
  int t = o.test_case;
  if (t == MHY_CASE_SOD_X || t == MHY_CASE_BI_SOD_X || t == MHY_CASE_SOD_Y || t == MHY_CASE_BI_SOD_Y ||
      t >= MHY_CASE_RIEMANN_PERTURBED) {
    leftBC = sym, leftBCValue = ey, rightBC = sym, rightBCValue = ey;
    topBC = sym, topBCValue = ex, bottomBC = sym, bottomBCValue = ex;
  }
  if (t == MHY_CASE_ADVECTION_X || t == MHY_CASE_BI_ADVECTION_X || t == MHY_CASE_ADVECTION_VIT_X ||
      t == MHY_CASE_BI_ADVECTION_VIT_X) {
    rightBC = imp, rightBCValue = {1.0, 0.0}, rightCellBC = MHY_BC_PERIODIC;
    leftBC = imp, leftBCValue = {1.0, 0.0};
    topBC = sym, topBCValue = ex, bottomBC = sym, bottomBCValue = ex;
  }
  if (t == MHY_CASE_ADVECTION_Y || t == MHY_CASE_BI_ADVECTION_Y) {
    bottomBC = imp, bottomBCValue = {0.0, 1.0}, bottomCellBC = MHY_BC_PERIODIC;
    topBC = imp, topBCValue = {0.0, 1.0}, topCellBC = MHY_BC_PERIODIC;
    leftBC = sym, leftBCValue = ey, rightBC = sym, rightBCValue = ey;
  }
  if (t == MHY_CASE_UNIT_TEST || t == MHY_CASE_BI_UNIT_TEST) {
    rightBC = imp, rightBCValue = {1.0, 0.0};
    leftBC = imp, leftBCValue = {1.0, 0.0};
    leftFluxBC = 1;
    leftFluxBCValue = ezero();
    leftFluxBCValue[0] = 1.0, leftFluxBCValue[1] = 1.0;
    bottomBC = imp, bottomBCValue = {1.0, 0.0};
    topBC = imp, topBCValue = {1.0, 0.0};
  }
  FluxBC = leftFluxBC + rightFluxBC + bottomFluxBC + topFluxBC;
}

// vnr_remap/Init.cc:495-511
void ref_ctx::initCellPos() {
  for (long c = 0; c < mesh.nbCells; ++c) {
    auto nd = mesh.nodesOfCell(c);
    V2 r = {0.0, 0.0};
    for (int s = 0; s < 4; ++s) r = vadd(r, node_coord_n0[nd[s]]);
    cell_coord_n0[c] = vmul(r, 0.25);
  }
}

// vnr_remap/Init.cc:103-394, plus the initialisers that are not in the
// reference (MHY_CASE_* >= 100).
int ref_ctx::init() {
  const long C = mesh.nbCells, N = mesh.nbNodes;
  const int t = o.test_case;
  for (long c = 0; c < C; ++c)
    for (int k = 0; k < NMAT; ++k) {
      fracvol_env[c][k] = 0.0, mass_fraction_env[c][k] = 0.0;
      density_env_n0[c][k] = 0.0, pressure_env_n0[c][k] = 0.0;
    }
  const double gamma = o.gamma;
  auto setVel = [&](long p, V2 u) {
    node_velocity_n0[p] = u;
    x_velocity[p] = u[0], y_velocity[p] = u[1];
  };
  if (t == MHY_CASE_SOD_X || t == MHY_CASE_SOD_Y) {
    for (long c = 0; c < C; ++c) {
      double r = (t == MHY_CASE_SOD_X) ? cell_coord_n0[c][0] : cell_coord_n0[c][1];
      fracvol_env[c][0] = 1.0, mass_fraction_env[c][0] = 1.0;
      if (r < 0.5) {
        density_n0[c] = 1.0, pressure_n0[c] = 1.0, density_env_n0[c][0] = 1.0, pressure_env_n0[c][0] = 1.0;
      } else {
        density_n0[c] = 0.125, pressure_n0[c] = 0.1, density_env_n0[c][0] = 0.125, pressure_env_n0[c][0] = 0.1;
      }
      // the reference's inverted sound speed sqrt(gamma*rho/p) (Init.cc:133-135)
      speed_env_n0[c][0] = std::sqrt(gamma * density_env_n0[c][0] / pressure_env_n0[c][0]);
      speed_n0[c] = speed_env_n0[c][0];
    }
    for (long p = 0; p < N; ++p) setVel(p, {0.0, 0.0});
  } else if (t == MHY_CASE_BI_SOD_X || t == MHY_CASE_BI_SOD_Y) {
    for (long c = 0; c < C; ++c) {
      double r = (t == MHY_CASE_BI_SOD_X) ? cell_coord_n0[c][0] : cell_coord_n0[c][1];
      if (r < 0.5) {
        fracvol_env[c][0] = 1.0, fracvol_env[c][1] = 0.0, mass_fraction_env[c][0] = 1.0;
        mass_fraction_env[c][1] = 0.0;
        density_n0[c] = 1.0, pressure_n0[c] = 1.0, density_env_n0[c][0] = 1.0, pressure_env_n0[c][0] = 1.0;
      } else {
        fracvol_env[c][0] = 0.0, fracvol_env[c][1] = 1.0, mass_fraction_env[c][0] = 0.0;
        mass_fraction_env[c][1] = 1.0;
        density_n0[c] = 0.125, pressure_n0[c] = 0.1, density_env_n0[c][1] = 0.125, pressure_env_n0[c][1] = 0.1;
      }
      speed_env_n0[c][0] = std::sqrt(gamma * density_env_n0[c][0] / pressure_env_n0[c][0]);
      speed_env_n0[c][1] = std::sqrt(gamma * density_env_n0[c][1] / pressure_env_n0[c][1]);
      speed_n0[c] = smin(speed_env_n0[c][0], speed_env_n0[c][1]);
    }
    for (long p = 0; p < N; ++p) setVel(p, {0.0, 0.0});
  } else if (t == MHY_CASE_ADVECTION_X || t == MHY_CASE_ADVECTION_Y || t == MHY_CASE_BI_ADVECTION_X ||
             t == MHY_CASE_BI_ADVECTION_Y) {
    const bool bi = (t == MHY_CASE_BI_ADVECTION_X || t == MHY_CASE_BI_ADVECTION_Y);
    const bool isx = (t == MHY_CASE_ADVECTION_X || t == MHY_CASE_BI_ADVECTION_X);
    for (long c = 0; c < C; ++c) {
      double r = isx ? cell_coord_n0[c][0] : cell_coord_n0[c][1];
      if (r < 0.3) {
        fracvol_env[c][0] = 1.0, mass_fraction_env[c][0] = 1.0;
        density_n0[c] = 1.0, pressure_n0[c] = 0.0, density_env_n0[c][0] = 1.0, pressure_env_n0[c][0] = 0.0;
      } else if ((r > 0.3) && (r < 0.5)) {
        if (!bi) {
          fracvol_env[c][0] = 1.0, mass_fraction_env[c][0] = 1.0;
          density_n0[c] = 10.0, pressure_n0[c] = 0.0, density_env_n0[c][0] = 10.0, pressure_env_n0[c][0] = 0.0;
        } else {
          fracvol_env[c][1] = 1.0, mass_fraction_env[c][1] = 1.0;
          density_n0[c] = 1.0, pressure_n0[c] = 0.0, density_env_n0[c][1] = 1.0, pressure_env_n0[c][1] = 0.0;
        }
      } else if (r > 0.5) {
        fracvol_env[c][0] = 1.0, mass_fraction_env[c][0] = 1.0;
        density_n0[c] = 1.0, pressure_n0[c] = 0.0, density_env_n0[c][0] = 1.0, pressure_env_n0[c][0] = 0.0;
      }
      speed_env_n0[c][0] = 1.0, speed_env_n0[c][1] = 1.0;
      speed_n0[c] = smin(speed_env_n0[c][0], speed_env_n0[c][1]);
      energy_env_n0[c][0] = 1.0, energy_env_n0[c][1] = 1.0, energy_n0[c] = 1.0;
    }
    V2 u = isx ? V2{1.0, 0.0} : V2{0.0, 1.0};
    for (long p = 0; p < N; ++p) setVel(p, u);
  } else if (t == MHY_CASE_BI_ADVECTION_VIT_X || t == MHY_CASE_BI_ADVECTION_VIT_Y) {
    const bool isx = (t == MHY_CASE_BI_ADVECTION_VIT_X);
    for (long c = 0; c < C; ++c) {
      double r = isx ? cell_coord_n0[c][0] : cell_coord_n0[c][1];
      if (r < 0.3) {
        fracvol_env[c][0] = 1.0, mass_fraction_env[c][0] = 1.0;
        density_n0[c] = 4.0, pressure_n0[c] = 0.0, density_env_n0[c][0] = 4.0, pressure_env_n0[c][0] = 0.0;
      } else if ((r > 0.3) && (r < 0.5)) {
        fracvol_env[c][1] = 1.0, mass_fraction_env[c][1] = 1.0;
        density_n0[c] = 4.0, pressure_n0[c] = 0.0, density_env_n0[c][1] = 1.0, pressure_env_n0[c][1] = 0.0;
      } else if (r > 0.5) {
        fracvol_env[c][0] = 1.0, mass_fraction_env[c][0] = 1.0;
        density_n0[c] = 4.0, pressure_n0[c] = 0.0, density_env_n0[c][0] = 4.0, pressure_env_n0[c][0] = 0.0;
      }
      speed_env_n0[c][0] = 1.0, speed_env_n0[c][1] = 1.0;
      speed_n0[c] = smin(speed_env_n0[c][0], speed_env_n0[c][1]);
      energy_env_n0[c][0] = 1.0, energy_env_n0[c][1] = 1.0, energy_n0[c] = 1.0;
    }
    V2 u = isx ? V2{1.0, 0.0} : V2{0.0, 1.0};
    for (long p = 0; p < N; ++p) {
      double r = isx ? node_coord_n0[p][0] : node_coord_n0[p][1];
      V2 v = u;
      if (r <= 0.3) v = vmul(u, 0.2);
      else if ((r > 0.3) && (r <= 0.5)) v = vmul(u, 4.0 * r - 1.0);
      setVel(p, v);
    }
  } else if (t == MHY_CASE_UNIT_TEST || t == MHY_CASE_BI_UNIT_TEST) {
    for (long c = 0; c < C; ++c) {
      fracvol_env[c][0] = 1.0, fracvol_env[c][1] = 0.0, mass_fraction_env[c][0] = 1.0;
      mass_fraction_env[c][1] = 0.0;
      pressure_n0[c] = 1.0, pressure_env_n0[c][0] = 1.0, pressure_env_n0[c][1] = 0.0;
      speed_n0[c] = std::sqrt(gamma), speed_env_n0[c][0] = std::sqrt(gamma), speed_env_n0[c][1] = std::sqrt(gamma);
      density_env_n0[c][1] = 0.0;
      double r = cell_coord_n0[c][0];
      if (r < 0.4) density_n0[c] = 1.0, density_env_n0[c][0] = 1.0;
      else if (r > 0.4) density_n0[c] = 0.1, density_env_n0[c][0] = 0.1;
      if (r > 0.6) density_n0[c] = 0.1, density_env_n0[c][0] = 0.1;
    }
    for (long p = 0; p < N; ++p) setVel(p, {1.0, 0.0});
  } else if (t >= MHY_CASE_RIEMANN_PERTURBED && t <= MHY_CASE_RIEMANN) {
    // Initialisers without a reference counterpart: pure or cut cells with
    // per-material states, physical sound speed sqrt(gamma_k p_k / rho_k).
    const int nm = o.nbmat;
    std::vector<double> rhok(NMAT, 0.0), pk(NMAT, 0.0);
    Rng rng(o.seed);
    // interface modes for MHY_CASE_RIEMANN_PERTURBED (drawn once)
    int nmodes = (int)o.param[2];
    std::vector<double> amp(nmodes > 0 ? nmodes : 1), ph(nmodes > 0 ? nmodes : 1);
    double asum = 0.0;
    if (t == MHY_CASE_RIEMANN_PERTURBED)
      for (int k = 0; k < nmodes; ++k) {
        amp[k] = 0.5 + 0.5 * rng.uniform();
        ph[k] = 2.0 * 3.14159265358979323846 * rng.uniform();
        asum += amp[k];
      }
    const double dx = o.dx, dy = o.dy;
    for (long c = 0; c < C; ++c) {
      long j = c / mesh.nx, i = c - j * mesh.nx;
      double f[NMAT] = {0.0, 0.0, 0.0};
      double rho[NMAT] = {0.0, 0.0, 0.0}, p[NMAT] = {0.0, 0.0, 0.0};
      if (t == MHY_CASE_RIEMANN_PERTURBED || t == MHY_CASE_RIEMANN) {
        double x0 = o.param[0];
        if (t == MHY_CASE_RIEMANN_PERTURBED) {
          rho[0] = o.param[4], p[0] = o.param[5], rho[1] = o.param[6], p[1] = o.param[7];
        } else {
          rho[0] = o.param[1], p[0] = o.param[2], rho[1] = o.param[3], p[1] = o.param[4];
        }
        double xl = node_coord_n0[mesh.node(j, i)][0];
        double cover = 0.0;
        const int ns = (t == MHY_CASE_RIEMANN_PERTURBED) ? 32 : 1;
        for (int s = 0; s < ns; ++s) {
          double ys = node_coord_n0[mesh.node(j, i)][1] + dy * ((double)s + 0.5) / (double)ns;
          double xi = x0;
          if (t == MHY_CASE_RIEMANN_PERTURBED) {
            double acc = 0.0;
            for (int k = 0; k < nmodes; ++k)
              acc += amp[k] * std::sin(2.0 * 3.14159265358979323846 * (double)(k + 1) * ys + ph[k]);
            xi = x0 + o.param[1] * acc / asum;
          }
          double w = (xi - xl) / dx;
          cover += (w < 0.0) ? 0.0 : (w > 1.0 ? 1.0 : w);
        }
        f[0] = cover / (double)ns;
        f[1] = 1.0 - f[0];
      } else if (t == MHY_CASE_POINT_EXPLOSION) {
        f[0] = 1.0, rho[0] = o.param[0], p[0] = o.param[1];
        if (j < 2 && i < 2) {
          double vol = dx * dy;
          double e = o.param[2] / 4.0 / (rho[0] * vol);
          p[0] = (o.gammap[0] - 1.0) * rho[0] * e;
        }
      } else {  // MHY_CASE_TRIPLE_LAYOUT
        double xc = cell_coord_n0[c][0], yc = cell_coord_n0[c][1];
        int k = (xc < 1.0) ? 0 : (yc >= 0.5 ? 1 : 2);
        const double r0[3] = {1.0, 0.125, 1.0}, p0[3] = {1.0, 0.1, 0.1};
        f[k] = 1.0;
        rho[k] = r0[k] * (1.0 + o.param[0] * (2.0 * rng.uniform() - 1.0));
        p[k] = p0[k];
      }
      double rmix = 0.0, pmix = 0.0;
      for (int k = 0; k < nm; ++k) {
        if (f[k] <= 0.0) rho[k] = 0.0, p[k] = 0.0;
        rmix += f[k] * rho[k];
        pmix += f[k] * p[k];
      }
      double cmax = 0.0;
      for (int k = 0; k < nm; ++k) {
        fracvol_env[c][k] = f[k];
        mass_fraction_env[c][k] = f[k] * rho[k] / rmix;
        density_env_n0[c][k] = rho[k];
        pressure_env_n0[c][k] = p[k];
        double ck = (f[k] > 0.0) ? std::sqrt(o.gammap[k] * p[k] / rho[k]) : 0.0;
        speed_env_n0[c][k] = ck;
        cmax = smax(cmax, ck);
      }
      density_n0[c] = rmix, pressure_n0[c] = pmix, speed_n0[c] = cmax;
    }
    for (long p = 0; p < N; ++p) setVel(p, {0.0, 0.0});
  } else {
    return MHY_ERR_UNKNOWN_CASE;
  }
  for (long c = 0; c < C; ++c) {
    fracvol_env1[c] = fracvol_env[c][0], fracvol_env2[c] = fracvol_env[c][1], fracvol_env3[c] = fracvol_env[c][2];
    int matcell = 0, imatpure = -1;
    for (int k = 0; k < NMAT; ++k)
      if (fracvol_env[c][k] > o.threshold) matcell++, imatpure = k;
    if (matcell > 1) mixte[c] = 1, pure[c] = -1;
    else mixte[c] = 0, pure[c] = imatpure;
  }
  return MHY_OK;
}

// vnr_remap/Init.cc:400-429
void ref_ctx::initSubVol() {
  for (long c = 0; c < mesh.nbCells; ++c) {
    auto nd = mesh.nodesOfCell(c);
    for (int s = 0; s < 4; ++s) {
      const V2& pm = node_coord_n0[nd[(s + 3) % 4]];
      const V2& pp = node_coord_n0[nd[(s + 1) % 4]];
      const V2 x1 = cell_coord_n0[c];
      const V2 x2 = vmul(vadd(pm, node_coord_n0[nd[s]]), 0.5);
      const V2 x3 = node_coord_n0[nd[s]];
      const V2 x4 = vmul(vadd(pp, node_coord_n0[nd[s]]), 0.5);
      node_cellvolume_n0[c][s] = 0.5 * (cross2(x1, x2) + cross2(x2, x3) + cross2(x3, x4) + cross2(x4, x1));
    }
    euler_volume[c] = o.dx * o.dy;
  }
}

// vnr_remap/Init.cc:517-554
void ref_ctx::initMeshGeometryForFaces() {
  const V2 ex = {1.0, 0.0}, ey = {0.0, 1.0};
  for (long f = 0; f < mesh.nbFaces; ++f) {
    long n1 = mesh.firstNode(f), n2 = mesh.secondNode(f);
    V2 X_face = vmul(vadd(node_coord_n0[n1], node_coord_n0[n2]), 0.5);
    V2 face_vec = vsub(node_coord_n0[n2], node_coord_n0[n1]);
    Xf[f] = X_face;
    faceLength[f] = norm2(face_vec);
    long b, fr;
    mesh.cellsOfFace(f, b, fr);
    for (long c : {b, fr}) {
      if (c < 0) continue;
      auto fc = mesh.facesOfCell(c);
      int slot = 0;
      while (fc[slot] != f) ++slot;
      V2 d = vsub(X_face, cell_coord_n0[c]);
      outerFaceNormal[c][slot] = vdiv(d, norm2(d));
    }
    faceNormal[f] = (std::fabs(dot2(face_vec, ex)) < o.threshold) ? ex : ey;
  }
}

// remap/UtilesMesh.cc:53-108 (serial face order; the two slots commute in
// every later use)
void ref_ctx::facesOfNode() {
  for (long p = 0; p < mesh.nbNodes; ++p) VerticalFaceOfNode[p] = {-1, -1}, HorizontalFaceOfNode[p] = {-1, -1};
  for (long f = 0; f < mesh.nbFaces; ++f) {
    if (!mesh.innerFace(f)) continue;
    auto& tab = mesh.isVertical(f) ? VerticalFaceOfNode : HorizontalFaceOfNode;
    for (long nd : {mesh.firstNode(f), mesh.secondNode(f)}) {
      if (tab[nd][0] == -1) tab[nd][0] = f;
      else tab[nd][1] = f;
    }
  }
}

// vnr_remap/PhaseLagrange.cc:14-24
void ref_ctx::computeCellMass() {
  for (long c = 0; c < mesh.nbCells; ++c) {
    cell_mass[c] = o.dx * o.dy * density_n0[c];
    for (int k = 0; k < o.nbmat; ++k) cell_mass_env[c][k] = mass_fraction_env[c][k] * cell_mass[c];
  }
}

// vnr_remap/Init.cc:435-457 (min reduction ignoring NaN values)
void ref_ctx::initDeltaT() {
  double red = std::numeric_limits<double>::max();
  for (long c = 0; c < mesh.nbCells; ++c) {
    double r1 = 0.0;
    for (int s = 0; s < 4; ++s) r1 = r1 + node_cellvolume_n0[c][s];
    red = smin(red, 0.1 * std::sqrt(r1) / speed_n0[c]);
  }
  o.deltat_init = red * 1.0E-6;
}

// vnr_remap/Init.cc:463-472
void ref_ctx::initInternalEnergy() {
  for (long c = 0; c < mesh.nbCells; ++c) {
    for (int k = 0; k < NMAT; ++k)
      energy_env_n0[c][k] = pressure_env_n0[c][k] / ((o.gammap[k] - 1.0) * density_env_n0[c][k]);
    if (o.test_case >= MHY_CASE_RIEMANN_PERTURBED) {
      // per-material gamma: mixture energy is the mass-weighted sum
      double e = 0.0;
      for (int k = 0; k < o.nbmat; ++k)
        if (fracvol_env[c][k] > 0.0) e += mass_fraction_env[c][k] * energy_env_n0[c][k];
      energy_n0[c] = e;
    } else {
      energy_n0[c] = pressure_n0[c] / ((o.gamma - 1.0) * density_n0[c]);
    }
  }
}

// vnr_remap/Init.cc:479-489
void ref_ctx::initPseudo() {
  for (long c = 0; c < mesh.nbCells; ++c) {
    tau_n0[c] = 1 / density_n0[c];
    divu_n0[c] = 0.0;
    pseudo_n0[c] = 0.0;
    for (int k = 0; k < NMAT; ++k) tau_env_n0[c][k] = 1 / density_env_n0[c][k], pseudo_env_n0[c][k] = 0.0;
  }
}

// vnr_remap/VnrRemap.cc:68-87
void ref_ctx::setUpTimeLoopN() {
  deltat_n = o.deltat_init;
  node_coord_n = node_coord_n0;
  node_cellvolume_n = node_cellvolume_n0;
  density_n = density_n0, density_env_n = density_env_n0;
  pressure_n = pressure_n0, pressure_env_n = pressure_env_n0;
  pseudo_n = pseudo_n0, pseudo_env_n = pseudo_env_n0;
  tau_n = tau_n0, tau_env_n = tau_env_n0;
  divu_n = divu_n0;
  speed_n = speed_n0, speed_env_n = speed_env_n0;
  energy_n = energy_n0, energy_env_n = energy_env_n0;
  node_velocity_n = node_velocity_n0;
  cell_coord_n = cell_coord_n0;
}

// vnr_remap/PhaseLagrange.cc:31-48 (mean, not quarter sum)
void ref_ctx::computeNodeMass() {
#pragma omp parallel for
  for (long p = 0; p < mesh.nbNodes; ++p) {
    long cells[4];
    int slot[4];
    int k = mesh.cellsOfNode(p, cells, slot);
    double r = 0.0;
    for (int q = 0; q < k; ++q) r = r + cell_mass[cells[q]];
    node_mass[p] = r / (double)(size_t)k;
  }
}

// vnr_remap/VnrRemap.cc:330-341
int ref_ctx::prologue() {
  if (euc()) return eucPrologue();
  initBoundaryConditions();
  initCellPos();
  int st = init();
  if (st != MHY_OK) return st;
  initSubVol();
  initMeshGeometryForFaces();
  facesOfNode();
  computeCellMass();
  if (o.sans_lagrange == 0) initDeltaT();
  if (o.sans_lagrange == 0) initInternalEnergy();
  if (o.sans_lagrange == 0) initPseudo();
  setUpTimeLoopN();
  computeNodeMass();
  return MHY_OK;
}

// ---------------------------------------------------------------- Lagrange

// vnr_remap/PhaseLagrange.cc:114-136
void ref_ctx::computeCornerNormal() {
#pragma omp parallel for
  for (long c = 0; c < mesh.nbCells; ++c) {
    auto nd = mesh.nodesOfCell(c);
    for (int s = 0; s < 4; ++s)
      cqs[c][s] = lpcNpc(node_coord_n[nd[s]], node_coord_n[nd[(s + 1) % 4]], node_coord_n[nd[(s + 3) % 4]]);
  }
}

// vnr_remap/VnrRemap.cc:14-50
void ref_ctx::computeDeltaT() {
  double Aveccfleuler = 0.0, cfl = 0.1;
  if (o.avec_projection == 1) Aveccfleuler = 1, cfl = 0.05;
  double red = std::numeric_limits<double>::max();
#pragma omp parallel
  {
    double acc = std::numeric_limits<double>::max();
#pragma omp for nowait
    for (long c = 0; c < mesh.nbCells; ++c) {
      auto nd = mesh.nodesOfCell(c);
      double uc = 0.0, r1 = 0.0;
      for (int s = 0; s < 4; ++s) {
        r1 = r1 + node_cellvolume_n[c][s];
        const V2& u = node_velocity_n[nd[s]];
        uc += std::sqrt(u[0] * u[0] + u[1] * u[1]) * 0.25;
      }
      acc = smin(acc, cfl * std::sqrt(r1) / (Aveccfleuler * uc + speed_n[c]));
    }
#pragma omp critical
    red = smin(red, acc);
  }
  deltat_nplus1 = smin(red, 1.05 * deltat_n);
}

// vnr_remap/PhaseLagrange.cc:167-197
void ref_ctx::updateVelocity() {
  const double dt = 0.5 * (deltat_nplus1 + deltat_n);
#pragma omp parallel for
  for (long p = 0; p < mesh.nbNodes; ++p) {
    if (!mesh.innerNode(p)) continue;
    long cells[4];
    int slot[4];
    int k = mesh.cellsOfNode(p, cells, slot);
    V2 r = {0.0, 0.0};
    for (int q = 0; q < k; ++q) {
      long c = cells[q];
      r = vadd(r, vmul(cqs[c][slot[q]], pressure_n[c] + pseudo_n[c]));
    }
    node_velocity_nplus1[p] = vadd(node_velocity_n[p], vmul(r, dt / node_mass[p]));
    x_velocity[p] = node_velocity_nplus1[p][0];
    y_velocity[p] = node_velocity_nplus1[p][1];
  }
}

// vnr_remap/ConditionsLimites.cc:12-185
void ref_ctx::updateVelocityBoundaryConditions() {
  const double dt = 0.5 * (deltat_nplus1 + deltat_n);
  const long nx = mesh.nx, ny = mesh.ny;
  auto side = [&](long p, int bc, const V2& axis, const V2& value) {
    if (bc == MHY_BC_SYMMETRY) {
      long cells[4];
      int slot[4];
      int k = mesh.cellsOfNode(p, cells, slot);
      V2 r = {0.0, 0.0};
      for (int q = 0; q < k; ++q) {
        long c = cells[q];
        const V2& C = cqs[c][slot[q]];
        r = vadd(r, vmul(vadd(C, symmetricVector(C, axis)), pressure_n[c] + pseudo_n[c]));
      }
      node_velocity_nplus1[p] = vadd(node_velocity_n[p], vmul(r, dt / (node_mass[p])));
    } else if (bc == MHY_BC_IMPOSED_VELOCITY) {
      node_velocity_nplus1[p] = value;
    }
  };
  const V2 ax = {1.0, 0.0}, ay = {0.0, 1.0};
  for (long i = 0; i <= nx; ++i) side(mesh.node(0, i), bottomBC, ax, bottomBCValue);
  for (long i = 0; i <= nx; ++i) side(mesh.node(ny, i), topBC, ax, topBCValue);
  for (long j = 0; j <= ny; ++j) side(mesh.node(j, 0), leftBC, ay, leftBCValue);
  for (long j = 0; j <= ny; ++j) side(mesh.node(j, nx), rightBC, ay, rightBCValue);
  auto corner = [&](long p, int bcA, int bcB, const V2& value) {
    if (bcA == MHY_BC_SYMMETRY && bcB == MHY_BC_SYMMETRY) node_velocity_nplus1[p] = {0.0, 0.0};
    else if (bcA == MHY_BC_IMPOSED_VELOCITY && bcB == MHY_BC_IMPOSED_VELOCITY) node_velocity_nplus1[p] = value;
  };
  corner(mesh.node(ny, 0), topBC, leftBC, leftBCValue);
  corner(mesh.node(ny, nx), topBC, rightBC, rightBCValue);
  corner(mesh.node(0, 0), bottomBC, leftBC, leftBCValue);
  corner(mesh.node(0, nx), bottomBC, rightBC, rightBCValue);
}

// vnr_remap/PhaseLagrange.cc:203-208
void ref_ctx::updatePosition() {
#pragma omp parallel for
  for (long p = 0; p < mesh.nbNodes; ++p)
    node_coord_nplus1[p] = vadd(node_coord_n[p], vmul(node_velocity_nplus1[p], deltat_nplus1));
}

// vnr_remap/PhaseLagrange.cc:214-230
void ref_ctx::updateCellPos() {
#pragma omp parallel for
  for (long c = 0; c < mesh.nbCells; ++c) {
    auto nd = mesh.nodesOfCell(c);
    V2 r = {0.0, 0.0};
    for (int s = 0; s < 4; ++s) r = vadd(r, node_coord_nplus1[nd[s]]);
    cell_coord_nplus1[c] = vmul(r, 0.25);
  }
}

// vnr_remap/PhaseLagrange.cc:236-264
void ref_ctx::computeSubVol() {
#pragma omp parallel for
  for (long c = 0; c < mesh.nbCells; ++c) {
    auto nd = mesh.nodesOfCell(c);
    for (int s = 0; s < 4; ++s) {
      const V2& xm = node_coord_nplus1[nd[(s + 3) % 4]];
      const V2& xp = node_coord_nplus1[nd[(s + 1) % 4]];
      const V2& x0 = node_coord_nplus1[nd[s]];
      const V2 x1 = cell_coord_nplus1[c];
      const V2 x2 = vmul(vadd(xm, x0), 0.5);
      const V2 x3 = x0;
      const V2 x4 = vmul(vadd(xp, x0), 0.5);
      node_cellvolume_nplus1[c][s] = 0.5 * (cross2(x1, x2) + cross2(x2, x3) + cross2(x3, x4) + cross2(x4, x1));
    }
  }
}

// vnr_remap/PhaseLagrange.cc:270-296
void ref_ctx::updateRho() {
#pragma omp parallel for
  for (long c = 0; c < mesh.nbCells; ++c) {
    double r = 0.0;
    for (int s = 0; s < 4; ++s) r = r + node_cellvolume_nplus1[c][s];
    vLagrange[c] = r;
    density_nplus1[c] = 0.0;
    for (int k = 0; k < o.nbmat; ++k) {
      if (fracvol_env[c][k] > o.threshold) density_env_nplus1[c][k] = cell_mass_env[c][k] / (fracvol_env[c][k] * r);
      density_nplus1[c] += fracvol_env[c][k] * density_env_nplus1[c][k];
    }
  }
}

// vnr_remap/ConditionsLimites.cc:186-209
void ref_ctx::updatePeriodicBoundaryConditions() {
  if (rightCellBC != MHY_BC_PERIODIC) return;
  for (long j = 0; j < mesh.ny; ++j) {
    long c = mesh.cell(j, mesh.nx - 1), L = mesh.cell(j, 0);
    density_nplus1[L] = density_n0[L];
    energy_nplus1[L] = energy_n0[L];
    density_nplus1[c] = density_nplus1[L];
    energy_nplus1[c] = energy_nplus1[L];
    for (int k = 0; k < o.nbmat; ++k) {
      density_env_nplus1[L][k] = density_env_n0[L][k];
      energy_env_nplus1[L][k] = energy_env_n0[L][k];
      density_env_nplus1[c][k] = density_env_nplus1[L][k];
      energy_env_nplus1[c][k] = energy_env_nplus1[L][k];
    }
  }
}

// vnr_remap/PhaseLagrange.cc:303-316
void ref_ctx::computeTau() {
#pragma omp parallel for
  for (long c = 0; c < mesh.nbCells; ++c) {
    tau_nplus1[c] = 0.5 * (1.0 / density_nplus1[c] + 1.0 / density_n[c]);
    for (int k = 0; k < o.nbmat; ++k) {
      tau_env_nplus1[c][k] = 0.0;
      if ((density_env_nplus1[c][k] > o.threshold) && (density_env_n[c][k] > o.threshold))
        tau_env_nplus1[c][k] = 0.5 * (1.0 / density_env_nplus1[c][k] + 1.0 / density_env_n[c][k]);
    }
  }
}

// vnr_remap/PhaseLagrange.cc:380-388
void ref_ctx::computeDivU() {
#pragma omp parallel for
  for (long c = 0; c < mesh.nbCells; ++c)
    divu_nplus1[c] = 1.0 / deltat_nplus1 * (1.0 / density_nplus1[c] - 1.0 / density_n[c]) / tau_nplus1[c];
}

// vnr_remap/PhaseLagrange.cc:54-107 (max reduction ignores NaN; the
// reference prints the capped cells)
void ref_ctx::computeArtificialViscosity() {
  double P = -std::numeric_limits<double>::max();
  for (long c = 0; c < mesh.nbCells; ++c)
    if (pressure_n[c] > P) P = pressure_n[c];
#pragma omp parallel for
  for (long c = 0; c < mesh.nbCells; ++c) {
    if (divu_n[c] < 0.0) {
      double r0 = 0.0, r1 = 0.0;
      for (int s = 0; s < 4; ++s) r0 = r0 + node_cellvolume_n[c][s];
      for (int s = 0; s < 4; ++s) r1 = r1 + node_cellvolume_n[c][s];
      pseudo_nplus1[c] = 1.0 / tau_nplus1[c] *
                         (-0.5 * std::sqrt(r0) * speed_n[c] * divu_nplus1[c] +
                          (o.gamma + 1) / 2.0 * r1 * divu_nplus1[c] * divu_nplus1[c]);
    } else {
      pseudo_nplus1[c] = 0.0;
    }
    if (pseudo_nplus1[c] > P) pseudo_nplus1[c] = P;
    for (int k = 0; k < o.nbmat; ++k) pseudo_env_nplus1[c][k] = fracvol_env[c][k] * pseudo_nplus1[c];
  }
}

// vnr_remap/PhaseLagrange.cc:324-373
int ref_ctx::updateEnergy() {
  int bad = 0;
#pragma omp parallel for reduction(| : bad)
  for (long c = 0; c < mesh.nbCells; ++c) {
    energy_nplus1[c] = 0.0;
    for (int k = 0; k < o.nbmat; ++k) {
      energy_env_nplus1[c][k] = 0.0;
      const double r1 = density_env_nplus1[c][k], r0 = density_env_n[c][k];
      if ((r1 > o.threshold) && (r0 > o.threshold)) {
        double pseudo = 0.0;
        const double q1 = pseudo_env_nplus1[c][k], q0 = pseudo_env_n[c][k];
        if ((o.pseudo_centree == 1) && ((q1 + q0) * (1.0 / r1 - 1.0 / r0) > 0.0)) pseudo = 0.5 * (q1 + q0);
        if (o.pseudo_centree == 0) pseudo = q1;
        const double den = 1 + 0.5 * (o.gammap[k] - 1.0) * r1 * (1.0 / r1 - 1.0 / r0);
        const double num = energy_env_n[c][k] - (0.5 * pressure_env_n[c][k] + pseudo) * (1.0 / r1 - 1.0 / r0);
        energy_env_nplus1[c][k] = num / den;
        energy_nplus1[c] += mass_fraction_env[c][k] * energy_env_nplus1[c][k];
        if (std::fabs(energy_env_nplus1[c][k]) > 5. && o.abort_on_energy) bad = 1;
      }
    }
  }
  return bad ? MHY_ERR_ENERGY : MHY_OK;
}

// vnr_remap/PhaseLagrange.cc:392-430 (perfect gas and void; the other
// equations of state only print "not yet programmed" in the reference)
void ref_ctx::computeEOS() {
  for (int k = 0; k < o.nbmat; ++k) {
    const double g = o.gammap[k];
    if (o.eos[k] == MHY_EOS_PERFECT_GAS) {
#pragma omp parallel for
      for (long c = 0; c < mesh.nbCells; ++c) {
        pressure_env_nplus1[c][k] = (g - 1.0) * density_env_nplus1[c][k] * energy_env_nplus1[c][k];
        speed_env_nplus1[c][k] = std::sqrt(g * (g - 1.0) * energy_env_nplus1[c][k]);
      }
    } else if (o.eos[k] == MHY_EOS_VOID) {
      for (long c = 0; c < mesh.nbCells; ++c) {
        pressure_env_nplus1[c][k] = pressure_env_n[c][k];
        speed_env_nplus1[c][k] = speed_env_n[c][k];
        energy_env_nplus1[c][k] = energy_env_n[c][k];
        energy_nplus1[c] = 0.0;
      }
    }
  }
}

// vnr_remap/PhaseLagrange.cc:464-485
void ref_ctx::computePressionMoyenne() {
#pragma omp parallel for
  for (long c = 0; c < mesh.nbCells; ++c) {
    pressure_nplus1[c] = 0.0;
    for (int k = 0; k < o.nbmat; ++k) {
      pressure_nplus1[c] += fracvol_env[c][k] * pressure_env_nplus1[c][k];
      speed_nplus1[c] = smax(speed_nplus1[c], speed_env_nplus1[c][k]);
    }
    if (density_nplus1[c] > 0.)
      speed_nplus1[c] = std::sqrt(o.gammap[0] * (o.gammap[0] - 1.0) * energy_nplus1[c]);
    for (int k = 0; k < o.nbmat; ++k)
      if (o.eos[k] == MHY_EOS_VOID) energy_nplus1[c] += mass_fraction_env[c][k] * energy_env_nplus1[c][k];
  }
}

// ------------------------------------------------------- remap preparation

// vnr_remap/PrepareRemap.cc:96-211
void ref_ctx::computeVariablesForRemap() {
  const int m = o.nbmat;
#pragma omp parallel for
  for (long c = 0; c < mesh.nbCells; ++c) {
    VE& U = ULagrange[c];
    const double V = vLagrange[c];
    for (int k = 0; k < m; ++k) {
      U[k] = fracvol_env[c][k] * V;
      U[m + k] = mass_fraction_env[c][k] * V * density_nplus1[c];
      U[2 * m + k] = mass_fraction_env[c][k] * V * density_nplus1[c] * energy_nplus1[c];
    }
    U[3 * m + 0] = 0.0, U[3 * m + 1] = 0.0, U[3 * m + 2] = 0.0;
    U[3 * m + 3] = pseudo_nplus1[c] * V;
    VE& P = Phi[c];
    if (o.pp == 1) {
      double somme_volume = 0.0;
      for (int k = 0; k < m; ++k) somme_volume += U[k];
      for (int k = 0; k < m; ++k) {
        P[k] = U[k] / somme_volume;
        P[m + k] = (U[k] != 0.) ? U[m + k] / U[k] : 0.0;
      }
      for (int k = 0; k < m; ++k) P[2 * m + k] = (U[m + k] != 0.) ? U[2 * m + k] / U[m + k] : 0.0;
      P[3 * m + 0] = 0.0, P[3 * m + 1] = 0.0, P[3 * m + 2] = 0.0;
      P[3 * m + 3] = pseudo_nplus1[c];
    } else {
      for (int e = 0; e < NEQ; ++e) P[e] = U[e] / V;
    }
    rLagrange[c] = density_nplus1[c];
  }
#pragma omp parallel for
  for (long p = 0; p < mesh.nbNodes; ++p) {
    const V2& u = node_velocity_nplus1[p];
    const double M = node_mass[p];
    XLagrange[p] = node_coord_nplus1[p];
    UDualLagrange[p][0] = M * u[0];
    UDualLagrange[p][1] = M * u[1];
    UDualLagrange[p][2] = M;
    if (o.projection_conservative == 1) UDualLagrange[p][3] = M * (u[0] * u[0] + u[1] * u[1]);
    DualPhi[p][0] = u[0];
    DualPhi[p][1] = u[1];
    DualPhi[p][2] = M;
    if (o.projection_conservative == 1) DualPhi[p][3] = (u[0] * u[0] + u[1] * u[1]);
  }
}

// vnr_remap/PrepareRemap.cc:65-90
void ref_ctx::computeCellQuantitesForRemap() {
#pragma omp parallel for
  for (long c = 0; c < mesh.nbCells; ++c) {
    auto nd = mesh.nodesOfCell(c);
    V2 r = {0.0, 0.0};
    for (int s = 0; s < 4; ++s) {
      const V2& a = XLagrange[nd[s]];
      const V2& b = XLagrange[nd[(s + 1) % 4]];
      r = vadd(r, vmul(vadd(a, b), cross2(a, b)));
    }
    XcLagrange[c] = vmul(r, 1.0 / (6.0 * vLagrange[c]));
  }
}

// vnr_remap/PrepareRemap.cc:14-64
void ref_ctx::computeFaceQuantitesForRemap() {
#pragma omp parallel for
  for (long f = 0; f < mesh.nbFaces; ++f) {
    if (mesh.innerFace(f)) {
      long b, fr;
      mesh.cellsOfFace(f, b, fr);
      deltaxLagrange[f] = dot2(vsub(XcLagrange[fr], XcLagrange[b]), faceNormal[f]);
    }
    long n1 = mesh.firstNode(f), n2 = mesh.secondNode(f);
    faceLengthLagrange[f] = norm2(vsub(XLagrange[n2], XLagrange[n1]));
    V2 r = {0.0, 0.0};
    r = vadd(r, node_velocity_nplus1[n1]);
    r = vadd(r, node_velocity_nplus1[n2]);
    faceNormalVelocity[f] = dot2(vmul(r, 0.5), faceNormal[f]);
  }
}

// ------------------------------------------------------------ remap helpers

// remap/UtilesRemap.cc:9-23
double ref_ctx::fluxLimiter(int lim, double r) const {
  if (lim == MHY_LIM_MINMOD) return smax(0.0, smin(1.0, r));
  if (lim == MHY_LIM_SUPERBEE) return smax(0.0, smax(smin(2.0 * r, 1.0), smin(r, 2.0)));
  if (lim == MHY_LIM_VANLEER) return (r <= 0.0) ? 0.0 : 2.0 * r / (1.0 + r);
  return 0.0;
}

// remap/UtilesRemap.cc:25-72
double ref_ctx::fluxLimiterPP(int lim, double gradplus, double gradmoins, double y0, double yplus, double ymoins,
                              double h0, double hplus, double hmoins) const {
  (void)y0;
  double grady = 0.0;
  if (gradplus * gradmoins < 0.0) return 0.;
  if (lim == MHY_LIM_MINMODG) {
    grady = ((yplus - ymoins) > 0.) ? smin(std::fabs(gradplus), std::fabs(gradmoins))
                                    : -smin(std::fabs(gradplus), std::fabs(gradmoins));
  } else if (lim == MHY_LIM_SUPERBEEG) {
    grady = ((yplus - ymoins) > 0.) ? smax(std::fabs(gradplus), std::fabs(gradmoins))
                                    : -smax(std::fabs(gradplus), std::fabs(gradmoins));
  } else if (lim == MHY_LIM_VANLEERG) {
    double lp = (h0 / 2. + hplus) / (h0 + hplus + hmoins);
    double lm = (h0 / 2. + hmoins) / (h0 + hplus + hmoins);
    if ((lp * gradplus + lm * gradmoins) != 0.) grady = gradplus * gradmoins / (lp * gradplus + lm * gradmoins);
    else grady = 0.;
  } else if (lim == MHY_LIM_ULTRABEEG) {
    grady = (yplus - ymoins) / h0;
  } else if (lim == MHY_LIM_ARITHMETICG) {
    double lp = (h0 / 2. + hplus) / (h0 + hplus + hmoins);
    double lm = (h0 / 2. + hmoins) / (h0 + hplus + hmoins);
    grady = lm * gradplus + lp * gradmoins;
  }
  double gMp = gradplus * (h0 + hplus) / h0;
  double gMm = gradmoins * (h0 + hmoins) / h0;
  double gM = smin(std::fabs(gMp), std::fabs(gMm));
  if ((yplus - ymoins) > 0.) grady = smin(std::fabs(gM), std::fabs(grady));
  else grady = -smin(std::fabs(gM), std::fabs(grady));
  return grady;
}

// remap/UtilesRemap.cc:74-125
double ref_ctx::computeY0(int lim, double y0, double yplus, double ymoins, double h0, double hplus, double hmoins,
                          int type) const {
  (void)y0;
  double y0plus = 0., y0moins = 0.;
  if (lim == MHY_LIM_MINMODG || lim == MHY_LIM_MINMOD) {
    y0plus = yplus, y0moins = ymoins;
  } else if (lim == MHY_LIM_SUPERBEEG || lim == MHY_LIM_SUPERBEE) {
    y0plus = ((h0 + hmoins) * yplus + h0 * ymoins) / (2 * h0 + hmoins);
    y0moins = ((h0 + hplus) * ymoins + h0 * yplus) / (2 * h0 + hplus);
  } else if (lim == MHY_LIM_VANLEERG || lim == MHY_LIM_VANLEER) {
    double a = smin(yplus, ymoins), b = smax(yplus, ymoins);
    double xplus = (h0 * h0 + 3 * h0 * hmoins + 2 * hmoins * hmoins) * yplus;
    double xmoins = (h0 * h0 + 3 * h0 * hplus + 2 * hplus * hplus) * ymoins;
    xplus += (h0 * h0 - h0 * hplus - 2 * hplus * hplus + 2 * h0 * hmoins) * ymoins;
    xmoins += (h0 * h0 - h0 * hmoins - 2 * hmoins * hmoins + 2 * h0 * hplus) * yplus;
    xplus /= (2 * h0 * h0 + 5 * h0 * hmoins + 2 * hmoins * hmoins - h0 * hplus - 2 * hplus * hplus);
    xmoins /= (2 * h0 * h0 + 5 * h0 * hplus + 2 * hplus * hplus - h0 * hmoins - 2 * hmoins * hmoins);
    y0plus = smin(smax(xplus, a), b);
    y0moins = smin(smax(xmoins, a), b);
  } else if (lim == MHY_LIM_ULTRABEEG) {
    y0plus = (yplus + ymoins) / 2.;
    y0moins = (yplus + ymoins) / 2.;
  } else if (lim == MHY_LIM_ARITHMETICG) {
    y0plus = ((h0 + hmoins + hplus) * yplus + h0 * ymoins) / (2 * h0 + hmoins + hplus);
    y0moins = ((h0 + hmoins + hplus) * ymoins + h0 * yplus) / (2 * h0 + hmoins + hplus);
  } else if (lim == 3000) {
    y0plus = yplus, y0moins = ymoins;
  }
  if (type == 0) return y0plus;
  if (type == 1) return y0moins;
  return 0.0;
}

// remap/UtilesRemap.cc:127-145
double ref_ctx::computexgxd(double y0, double yplus, double ymoins, double h0, double y0plus, double y0moins,
                            int type) const {
  double xplus = 1.;
  if (std::fabs(y0plus - yplus) > o.threshold) xplus = (y0 - yplus) / (y0plus - yplus) - 1. / 2.;
  double xmoins = 1.;
  if (std::fabs(y0moins - ymoins) > o.threshold) xmoins = (y0 - ymoins) / (y0moins - ymoins) - 1. / 2.;
  double xd = +h0 * smin(smax(xplus, -1. / 2.), 1. / 2.);
  double xg = -h0 * smin(smax(xmoins, -1. / 2.), 1. / 2.);
  if (type == 0) return xg;
  if (type == 1) return xd;
  return 0.0;
}

// remap/UtilesRemap.cc:147-166
double ref_ctx::computeygyd(double y0, double yplus, double ymoins, double h0, double y0plus, double y0moins,
                            double grady, int type) const {
  double xtd = y0 + h0 / 2 * grady;
  double xtg = y0 - h0 / 2 * grady;
  double ad = smin(yplus, 2. * y0moins - ymoins);
  double bd = smax(yplus, 2. * y0moins - ymoins);
  double ag = smin(ymoins, 2. * y0plus - yplus);
  double bg = smax(ymoins, 2. * y0plus - yplus);
  double yd = smin(smax(xtd, ad), bd);
  double yg = smin(smax(xtg, ag), bg);
  if (type == 0) return yg;
  if (type == 1) return yd;
  return 0.0;
}

// remap/UtilesRemap.cc:177-188: integral of the segment (x0,y0)-(x1,y1)
// from x0 to X (clamped)
double ref_ctx::INT2Y(double X, double x0, double y0, double x1, double y1) {
  double flux = 0.;
  if (std::fabs(x1 - x0) > 1.e-14) {
    double eta = smin(smax(0., (X - x0) / (x1 - x0)), 1.);
    flux = (y0 + 0.5 * eta * (y1 - y0)) * (x1 - x0) * eta;
  }
  return flux;
}

// remap/UtilesRemap-Impl.h:99-126
VE ref_ctx::computeAndLimitGradPhi(int lim, const VE& gp, const VE& gm, const VE& phi, const VE& phip,
                                   const VE& phim, double h0, double hp, double hm) const {
  VE res;
  if (lim < MHY_LIM_MINMODG) {
    for (int i = 0; i < NEQ; ++i)
      res[i] = (fluxLimiter(lim, divideNoExcept(gp[i], gm[i])) * gm[i] +
                fluxLimiter(lim, divideNoExcept(gm[i], gp[i])) * gp[i]) /
               2.0;
  } else {
    for (int i = 0; i < NEQ; ++i) res[i] = fluxLimiterPP(lim, gp[i], gm[i], phi[i], phip[i], phim[i], h0, hp, hm);
  }
  return res;
}

// remap/UtilesRemap-Impl.h:142-489 (pureVariant selects computeFluxPPPure)
void ref_ctx::computeFluxPP(bool pureVariant, const VE& gradphi, const VE& phi, const VE& phiplus,
                            const VE& phimoins, double h0, double hplus, double hmoins, double fnv, double dt,
                            int type, double thr, int complet, double dualv, int calcDual, VE* pFlux,
                            VE* pFluxDual) const {
  VE Flux = ezero(), Flux_dual = ezero();
  const int m = o.nbmat;
  const int lim = pureVariant ? o.limiter_pure : o.limiter;
  const double ppv = 0.5 * (fnv + std::fabs(fnv)) * dt;
  const double ppdv = 0.5 * (dualv + std::fabs(dualv)) * dt;
  for (int i = 0; i < NEQ; ++i) {
    double y0plus = computeY0(lim, phi[i], phiplus[i], phimoins[i], h0, hplus, hmoins, 0);
    double y0moins = computeY0(lim, phi[i], phiplus[i], phimoins[i], h0, hplus, hmoins, 1);
    double xg = computexgxd(phi[i], phiplus[i], phimoins[i], h0, y0plus, y0moins, 0);
    double xd = computexgxd(phi[i], phiplus[i], phimoins[i], h0, y0plus, y0moins, 1);
    double yg = computeygyd(phi[i], phiplus[i], phimoins[i], h0, y0plus, y0moins, gradphi[i], 0);
    double yd = computeygyd(phi[i], phiplus[i], phimoins[i], h0, y0plus, y0moins, gradphi[i], 1);
    double f1, f2, f3, f1m, f2m, f3m;
    const bool flat = ((phiplus[i] - phi[i]) * (phimoins[i] - phi[i])) >= 0.;
    if (type == 0) {
      f1m = INT2Y(-h0 / 2. + ppv, -h0 / 2., phimoins[i], xg, yg);
      f1 = INT2Y(-h0 / 2., -h0 / 2., phimoins[i], xg, yg);
      f2m = INT2Y(-h0 / 2. + ppv, xg, yg, xd, yd);
      f2 = INT2Y(-h0 / 2., xg, yg, xd, yd);
      f3m = INT2Y(-h0 / 2. + ppv, xd, yd, h0 / 2., phiplus[i]);
      f3 = INT2Y(-h0 / 2., xd, yd, h0 / 2., phiplus[i]);
      Flux[i] = smax(((f1m - f1) + (f2m - f2) + (f3m - f3)), 0.);
      if (flat) Flux[i] = phi[i] * ppv;
      if (calcDual == 1) {
        f1m = INT2Y(ppdv, -h0 / 2., phimoins[i], xg, yg);
        f1 = INT2Y(0., -h0 / 2., phimoins[i], xg, yg);
        f2m = INT2Y(ppdv, xg, yg, xd, yd);
        f2 = INT2Y(0., xg, yg, xd, yd);
        f3m = INT2Y(ppdv, xd, yd, h0 / 2., phiplus[i]);
        f3 = INT2Y(0., xd, yd, h0 / 2., phiplus[i]);
        Flux_dual[i] = smax(((f1m - f1) + (f2m - f2) + (f3m - f3)), 0.);
        if (flat) Flux_dual[i] = phi[i] * ppdv;
      }
    } else {
      f1 = INT2Y(h0 / 2. - ppv, -h0 / 2., phimoins[i], xg, yg);
      f1m = INT2Y(h0 / 2., -h0 / 2., phimoins[i], xg, yg);
      f2 = INT2Y(h0 / 2. - ppv, xg, yg, xd, yd);
      f2m = INT2Y(h0 / 2., xg, yg, xd, yd);
      f3 = INT2Y(h0 / 2. - ppv, xd, yd, h0 / 2., phiplus[i]);
      f3m = INT2Y(h0 / 2., xd, yd, h0 / 2., phiplus[i]);
      Flux[i] = smax(((f1m - f1) + (f2m - f2) + (f3m - f3)), 0.);
      if (flat) Flux[i] = phi[i] * ppv;
      if (calcDual == 1) Flux_dual[i] = 0.;
    }
  }
  if (complet == 1) {
    double sfm = 0., sfv = 0.;
    for (int k = 0; k < m; ++k) sfv += Flux[k];
    if (std::fabs(sfv) > thr) {
      for (int k = 0; k < m; ++k) {
        if (!pureVariant) Flux[m + k] = (Flux[m + k] / sfv) * Flux[k];
        Flux[2 * m + k] = (Flux[2 * m + k] / sfv) * Flux[m + k];
        sfm += Flux[m + k];
      }
      Flux[3 * m] = (Flux[3 * m] / sfv) * sfm;
      Flux[3 * m + 1] = (Flux[3 * m + 1] / sfv) * sfm;
      Flux[3 * m + 2] = (Flux[3 * m + 2] / sfv) * sfm;
      Flux[3 * m + 3] = phi[3 * m + 3] * sfv;
    } else {
      Flux = ezero();
    }
  } else {
    double sfm = 0., sfv = 0.;
    for (int k = 0; k < m; ++k) {
      if (!pureVariant) Flux[m + k] = phi[m + k] * Flux[k];
      Flux[2 * m + k] = phi[2 * m + k] * Flux[m + k];
      sfm += Flux[m + k];
      sfv += Flux[k];
    }
    Flux[3 * m] = phi[3 * m] * sfm;
    Flux[3 * m + 1] = phi[3 * m + 1] * sfm;
    Flux[3 * m + 2] = phi[3 * m + 2] * sfm;
    Flux[3 * m + 3] = phi[3 * m + 3] * sfv;
  }
  *pFlux = Flux;
  *pFluxDual = Flux_dual;
  if (ppv == 0.) *pFlux = ezero();
  if (ppdv == 0.) *pFluxDual = ezero();
}

// remap/UtilesRemap-Impl.h:503-524
VE ref_ctx::computeRemapFlux(double fnv, const V2& fn, double fl, const VE& phiFace, const V2& ofn, const V2& exy,
                             double dt) const {
  VE F;
  if (std::fabs(dot2(fn, exy)) < 1.0E-10) {
    for (int i = 0; i < NEQ; ++i) F[i] = phiFace[i] * 0.0;
    return F;
  }
  double s = (o.pp == 0) ? dot2(ofn, exy) * fnv * fl * dt : dot2(ofn, exy) * fl;
  for (int i = 0; i < NEQ; ++i) F[i] = phiFace[i] * s;
  return F;
}

// remap/BoundaryFluxes.cc:14-120 (inflow through the domain edge; the
// top-cell test keeps the reference's bound cCells <= nx*(ny-1))
VE ref_ctx::computeBoundaryFluxes(int proj, long c, const V2& exy) const {
  const long nx = mesh.nx, ny = mesh.ny;
  const auto& phiFace = (proj == 1) ? phiFace1 : phiFace2;
  auto fc = mesh.facesOfCell(c);
  if (bottomFluxBC == 1 && c < nx && exy[1] == 1)
    return computeRemapFlux(faceNormalVelocity[fc[0]], faceNormal[fc[0]], faceLength[fc[0]], phiFace[fc[3]],
                            outerFaceNormal[c][0], exy, deltat_n);
  if (topFluxBC == 1 && c <= nx * (ny - 1) && c < nx * ny && exy[1] == 1)
    return computeRemapFlux(faceNormalVelocity[fc[3]], faceNormal[fc[3]], faceLength[fc[3]], phiFace[fc[0]],
                            outerFaceNormal[c][3], exy, deltat_n);
  if (leftFluxBC == 1 && exy[0] == 1 && c % nx == 0)
    return computeRemapFlux(faceNormalVelocity[fc[1]], faceNormal[fc[1]], faceLength[fc[1]], phiFace[fc[2]],
                            outerFaceNormal[c][1], exy, deltat_n);
  if (rightFluxBC == 1 && exy[0] == 1 && c % nx == nx - 1)
    return computeRemapFlux(faceNormalVelocity[fc[2]], faceNormal[fc[2]], faceLength[fc[2]], phiFace[fc[1]],
                            outerFaceNormal[c][2], exy, deltat_n);
  return ezero();
}

// remap/UtilesRemap.cc:200-320 (third-order flux with TVD clamp)
double ref_ctx::ComputeFluxOrdre3(double ymmm, double ymm, double ym, double yp, double ypp, double yppp,
                                  double hmmm, double hmm, double hm, double hp, double hpp, double hppp,
                                  double vdt) const {
  auto grad = [](double hm_, double h0_, double hp_, double ym_, double y0_, double yp_) {
    return h0_ / (hm_ + h0_ + hp_) *
           ((2. * hm_ + h0_) / (h0_ + hp_) * (yp_ - y0_) + (h0_ + 2. * hp_) / (hm_ + h0_) * (y0_ - ym_));
  };
  auto ystar = [](double hmm_, double hm_, double hp_, double hpp_, double ymm_, double ym_, double yp_,
                  double ypp_, double gm_, double gp_) {
    double t1 = (2. * hp_ * hm_) / (hm_ + hp_) * ((hmm_ + hm_) / (2. * hm_ + hp_) - (hpp_ + hp_) / (2. * hp_ + hm_)) *
                (yp_ - ym_);
    double t2 = -hm_ * (hmm_ + hm_) / (2. * hm_ + hp_) * gp_ + hp_ * (hp_ + hpp_) / (hm_ + 2. * hp_) * gm_;
    return ym_ + hm_ / (hm_ + hp_) * (yp_ - ym_) + 1. / (hmm_ + hm_ + hp_ + hpp_) * (t1 + t2);
  };
  double gmm = grad(hmmm, hmm, hm, ymmm, ymm, ym);
  double gm = grad(hmm, hm, hp, ymm, ym, yp);
  double gp = grad(hm, hp, hpp, ym, yp, ypp);
  double gpp = grad(hp, hpp, hppp, yp, ypp, yppp);
  double ysm = ystar(hmmm, hmm, hm, hp, ymmm, ymm, ym, yp, gmm, gm);
  double ys = ystar(hmm, hm, hp, hpp, ymm, ym, yp, ypp, gm, gp);
  double ysp = ystar(hm, hp, hpp, hppp, ym, yp, ypp, yppp, gp, gpp);
  double ym_m = ysm, ym_p = ys, yp_m = ys, yp_p = ysp;
  double grad_m = ym_p - ym_m, grad_p = yp_p - yp_m;
  double ym6 = 6. * (ym - 0.5 * (ym_m + ym_p));
  double yp6 = 6. * (yp - 0.5 * (yp_m + yp_p));
  double flux;
  if (vdt >= 0.) {
    double x = vdt, dx = hm;
    flux = ym_p - 0.5 * x / dx * (grad_m - (1. - 2. / 3. * x / dx) * ym6);
  } else {
    double x = -vdt, dx = hp;
    flux = yp_m + 0.5 * x / dx * (grad_p - (1. - 2. / 3. * x / dx) * yp6);
  }
  double num = vdt / hm, nup = vdt / hp;
  double ym_ym = ym + (1. - num) / num * (ym - ymm);
  double yp_ym = yp - (1. + nup) / nup * (yp - ypp);
  double i1inf, i1sup, i2inf, i2sup;
  i1inf = smin(ym, yp), i1sup = smax(ym, yp);
  if (vdt >= 0.) i2inf = smin(ym, ym_ym), i2sup = smax(ym, ym_ym);
  else i2inf = smin(yp, yp_ym), i2sup = smax(yp, yp_ym);
  double linf, lsup;
  if ((i1sup < i2inf) || (i2sup < i1inf)) linf = 0., lsup = 0.;
  else linf = smax(i1inf, i2inf), lsup = smin(i1sup, i2sup);
  if (flux < linf) flux = linf;
  if (flux > lsup) flux = lsup;
  return flux;
}

// ----------------------------------------------------------- remap sweeps

// remap/PhaseRemap1.cc:22-122, remap/PhaseRemap2.cc:21-119
void ref_ctx::computeGradPhiFace(int sweep) {
  if (o.projection_order <= 1) return;
  const bool xdir = (sweep == 1) ? x_then_y_n : !x_then_y_n;
  auto& gPF = (sweep == 1) ? gradPhiFace1 : gradPhiFace2;
#pragma omp parallel for
  for (long f = 0; f < mesh.nbFaces; ++f) {
    if (!mesh.innerFace(f) || mesh.isVertical(f) != xdir) continue;
    long b, fr;
    mesh.cellsOfFace(f, b, fr);
    for (int e = 0; e < NEQ; ++e) gPF[f][e] = (Phi[fr][e] - Phi[b][e]) / deltaxLagrange[f];
  }
#pragma omp parallel for
  for (long c = 0; c < mesh.nbCells; ++c) {
    auto fc = mesh.facesOfCell(c);
    HvLagrange[c] = 0.;
    if (xdir) {
      HvLagrange[c] += 0.5 * faceLengthLagrange[fc[0]];
      HvLagrange[c] += 0.5 * faceLengthLagrange[fc[3]];
    } else {
      HvLagrange[c] += 0.5 * faceLengthLagrange[fc[2]];
      HvLagrange[c] += 0.5 * faceLengthLagrange[fc[1]];
    }
  }
}

// remap/PhaseRemap1.cc:128-323, remap/PhaseRemap2.cc:125-318
void ref_ctx::computeGradPhi(int sweep) {
  if (o.projection_order <= 1) return;
  const bool xdir = (sweep == 1) ? x_then_y_n : !x_then_y_n;
  auto& gPF = (sweep == 1) ? gradPhiFace1 : gradPhiFace2;
  auto& gP = (sweep == 1) ? gradPhi1 : gradPhi2;
  auto& DPF = (sweep == 1) ? DualphiFlux1 : DualphiFlux2;
#pragma omp parallel for
  for (long c = 0; c < mesh.nbCells; ++c) {
    auto fc = mesh.facesOfCell(c);
    // minus face / plus face and the neighbours through them
    long fm, fp, cm, cp;
    int sm, sp;
    if (xdir) {
      fp = fc[2], fm = fc[1], sp = 2, sm = 1;
      cp = mesh.frontCell(fp), cm = mesh.backCell(fm);
    } else {
      fm = fc[0], fp = fc[3], sm = 0, sp = 3;
      cm = mesh.frontCell(fm), cp = mesh.backCell(fp);
    }
    if (cm == -1) cm = c;
    if (cp == -1) cp = c;
    const int mix3[3] = {mixte[c], mixte[cp], mixte[cm]}, pure3[3] = {pure[c], pure[cp], pure[cm]};
    // x: (plus face, minus face) gradients; y: (minus, plus), as the reference passes them
    const VE& ga = xdir ? gPF[fp] : gPF[fm];
    const VE& gb = xdir ? gPF[fm] : gPF[fp];
    double Fm = 0.0, Fp = 0.0, fdual = 0.0;
    if (o.pp == 1) {
      const V2 exy = xdir ? V2{1.0, 0.0} : V2{0.0, 1.0};
      Fm = dot2(outerFaceNormal[c][sm], exy) * faceNormalVelocity[fm];
      Fp = dot2(outerFaceNormal[c][sp], exy) * faceNormalVelocity[fp];
      fdual = 0.5 * (faceNormalVelocity[fm] + faceNormalVelocity[fp]);
    }
    gradPhiCell(ga, gb, Phi[c], Phi[cp], Phi[cm], mix3, pure3, HvLagrange[c], HvLagrange[cp], HvLagrange[cm], Fm,
                Fp, fdual, deltat_n, &gP[c], &deltaPhiFaceAr[c], &deltaPhiFaceAv[c], &DPF[c]);
    if (o.pp == 1) {
      const double w = (faceLength[fm] + faceLength[fp]) * 0.5;
      for (int e = 0; e < NEQ; ++e) DPF[c][e] *= w;
    }
  }
}

// The cell stage of computeGradPhi{1,2} (PhaseRemap1.cc:134-320,
// PhaseRemap2.cc:126-318) for one cell c with its plus / minus neighbours:
// the PenteBorne-mixte test of a pure neighbourhood (c and both neighbours
// not mixed and holding the same material; flags in the order c, plus,
// minus), the limiter it selects, the limited gradient, and with PenteBorne
// the back (Ar, type 0) and front (Av, type 1) swept fluxes of
// computeFluxPP or, in a pure neighbourhood, computeFluxPPPure, plus the PB
// dual flux (before its face-length weight).
void ref_ctx::gradPhiCell(const VE& gp, const VE& gm, const VE& phi, const VE& phip, const VE& phim,
                          const int* mix3, const int* pure3, double h0, double hp, double hm, double Fm, double Fp,
                          double fdual, double dt, VE* grad, VE* Ar, VE* Av, VE* dual) const {
  const bool vp = (o.pp_mixte == 1) &&
                  (mix3[0] == 0 && mix3[1] == 0 && mix3[2] == 0 && pure3[0] == pure3[1] && pure3[0] == pure3[2]);
  int lim = o.limiter;
  if ((o.pp == 1) && vp) lim = o.limiter_pure;
  *grad = computeAndLimitGradPhi(lim, gp, gm, phi, phip, phim, h0, hp, hm);
  if (o.pp == 1) {
    const int calcDual = (o.methode_flux_masse == 2) ? 1 : 0;
    VE bidon;
    computeFluxPP(vp, *grad, phi, phip, phim, h0, hp, hm, Fm, dt, 0, o.threshold, o.pp_complet, fdual, calcDual, Ar,
                  dual);
    computeFluxPP(vp, *grad, phi, phip, phim, h0, hp, hm, Fp, dt, 1, o.threshold, o.pp_complet, fdual, calcDual, Av,
                  &bidon);
  }
}

// remap/PhaseRemap1.cc:331-452, remap/PhaseRemap2.cc:326-447
void ref_ctx::computeUpwindFaceQuantitiesForProjection(int sweep) {
  const bool xdir = (sweep == 1) ? x_then_y_n : !x_then_y_n;
  const auto& Uin = (sweep == 1) ? ULagrange : Uremap1;
  auto& gP = (sweep == 1) ? gradPhi1 : gradPhi2;
  auto& pF = (sweep == 1) ? phiFace1 : phiFace2;
#pragma omp parallel for
  for (long f = 0; f < mesh.nbFaces; ++f) {
    if (!mesh.innerFace(f) || mesh.isVertical(f) != xdir) continue;
    long cb, cf;
    mesh.cellsOfFace(f, cb, cf);
    if (o.projection_order <= 2) {
      if (o.pp == 0) {
        const double vel = faceNormalVelocity[f], dx = deltaxLagrange[f];
        long up = (vel * dx > 0.0) ? cb : cf;
        const double s = dot2(vsub(Xf[f], XcLagrange[up]), faceNormal[f]);
        for (int e = 0; e < NEQ; ++e) pF[f][e] = Uin[up][e] / vLagrange[up] + gP[up][e] * s;
      } else {
        // X: Av(back = left) - Ar(front = right); Y: Av(front = below) - Ar(back = above)
        for (int e = 0; e < NEQ; ++e)
          pF[f][e] = xdir ? deltaPhiFaceAv[cb][e] - deltaPhiFaceAr[cf][e] : deltaPhiFaceAv[cf][e] - deltaPhiFaceAr[cb][e];
      }
    } else if (o.projection_order == 3) {
      long c0, c1, c2, d0, d1, d2;  // upstream side (minus) and downstream side (plus)
      if (xdir) {
        c0 = cb, c1 = mesh.leftCells(c0), c2 = mesh.leftCells(c1);
        d0 = cf, d1 = mesh.rightCells(d0), d2 = mesh.rightCells(d1);
      } else {
        c0 = cf, c1 = mesh.bottomCells(c0), c2 = mesh.bottomCells(c1);
        d0 = cb, d1 = mesh.topCells(d0), d2 = mesh.topCells(d1);
      }
      for (int e = 0; e < NEQ; ++e)
        pF[f][e] = ComputeFluxOrdre3(Uin[c2][e] / vLagrange[c2], Uin[c1][e] / vLagrange[c1], Uin[c0][e] / vLagrange[c0],
                                     Uin[d0][e] / vLagrange[d0], Uin[d1][e] / vLagrange[d1], Uin[d2][e] / vLagrange[d2],
                                     HvLagrange[c2], HvLagrange[c1], HvLagrange[c0], HvLagrange[d0], HvLagrange[d1],
                                     HvLagrange[d2], faceNormalVelocity[f] * deltat_n);
    }
  }
}

// remap/PhaseRemap1.cc:459-568, remap/PhaseRemap2.cc:454-497
void ref_ctx::computeUremap(int sweep) {
  const int m = o.nbmat;
  const bool xdir = (sweep == 1) ? x_then_y_n : !x_then_y_n;
  const V2 exy = xdir ? V2{1.0, 0.0} : V2{0.0, 1.0};
  const auto& Uin = (sweep == 1) ? ULagrange : Uremap1;
  auto& Uout = (sweep == 1) ? Uremap1 : Uremap2;
  auto& pF = (sweep == 1) ? phiFace1 : phiFace2;
  auto& FF = (sweep == 1) ? FluxFace1 : FluxFace2;
#pragma omp parallel for
  for (long c = 0; c < mesh.nbCells; ++c) {
    long j = c / mesh.nx, i = c - j * mesh.nx;
    auto fc = mesh.facesOfCell(c);
    VE red = ezero();
    // neighbour order below, above, left, right (CartesianMesh2D.cc:155-166)
    const int slots[4] = {0, 3, 1, 2};
    const bool exists[4] = {j >= 1, j < mesh.ny - 1, i >= 1, i < mesh.nx - 1};
    for (int q = 0; q < 4; ++q) {
      if (!exists[q]) continue;
      int s = slots[q];
      long f = fc[s];
      VE F = computeRemapFlux(faceNormalVelocity[f], faceNormal[f], faceLength[f], pF[f], outerFaceNormal[c][s], exy,
                              deltat_n);
      FF[c][s] = F;
      for (int e = 0; e < NEQ; ++e) red[e] = red[e] + F[e];
    }
    if (FluxBC > 0) {
      VE B = computeBoundaryFluxes(sweep, c, exy);
      for (int e = 0; e < NEQ; ++e) red[e] = red[e] + B[e];
    }
    for (int e = 0; e < NEQ; ++e) Uout[c][e] = Uin[c][e] - red[e];
    if (sweep != 1) continue;
    const VE& U1 = Uout[c];
    VE& P = Phi[c];
    if (o.pp == 1) {
      double sv = 0.0, sm = 0.0;
      for (int k = 0; k < m; ++k) sv += U1[k];
      for (int k = 0; k < m; ++k) {
        P[k] = U1[k] / sv;
        P[m + k] = (U1[k] != 0.) ? U1[m + k] / (U1[k]) : 0.0;
        sm += U1[m + k];
      }
      P[3 * m] = U1[3 * m] / sm;
      P[3 * m + 1] = U1[3 * m + 1] / sm;
      for (int k = 0; k < m; ++k) P[2 * m + k] = (U1[m + k] != 0.) ? U1[2 * m + k] / U1[m + k] : 0.0;
      if (o.projection_conservative == 1) P[3 * m + 2] = U1[3 * m + 2] / sm;
    } else {
      for (int e = 0; e < NEQ; ++e) P[e] = U1[e] / vLagrange[c];
    }
    int matcell = 0, imatpure = -1;
    for (int k = 0; k < m; ++k)
      if (P[k] > 0.) matcell++, imatpure = k;
    if (matcell > 1) mixte[c] = 1, pure[c] = -1;
    else mixte[c] = 0, pure[c] = imatpure;
  }
}

// remap/UtilesRemap-Impl.h:8-97 (horizontal: along x with left/right nodes)
VE ref_ctx::dualGradPhi(bool horizontal, long p) const {
  long L = horizontal ? mesh.leftNode(p) : mesh.bottomNode(p);
  long R = horizontal ? mesh.rightNode(p) : mesh.topNode(p);
  if (L == -1 || R == -1) return ezero();
  const int a = horizontal ? 0 : 1;
  VE gr = ezero(), gl = ezero();
  gr[0] = (DualPhi[R][0] - DualPhi[p][0]) / (XLagrange[R][a] - XLagrange[p][a]);
  gl[0] = (DualPhi[p][0] - DualPhi[L][0]) / (XLagrange[p][a] - XLagrange[L][a]);
  gr[1] = (DualPhi[R][1] - DualPhi[p][1]) / (XLagrange[R][a] - XLagrange[p][a]);
  gl[1] = (DualPhi[p][1] - DualPhi[L][1]) / (XLagrange[p][a] - XLagrange[L][a]);
  long LL = horizontal ? mesh.leftNode(L) : mesh.bottomNode(L);
  long RR = horizontal ? mesh.rightNode(R) : mesh.topNode(R);
  // the reference reads index -1 when both outer neighbours are missing;
  // the widths only enter the generalised limiters
  auto X = [&](long q) { return q >= 0 ? XLagrange[q][a] : 0.0; };
  double h0 = 0.5 * (XLagrange[R][a] - XLagrange[L][a]), hm, hp;
  if (LL == -1) {
    hm = 0., hp = 0.5 * (X(RR) - XLagrange[p][a]);
  } else if (RR == -1) {
    hp = 0., hm = 0.5 * (XLagrange[p][a] - X(LL));
  } else {
    hm = 0.5 * (XLagrange[p][a] - XLagrange[LL][a]);
    hp = 0.5 * (XLagrange[RR][a] - XLagrange[p][a]);
  }
  return computeAndLimitGradPhi(o.limiter, gr, gl, DualPhi[p], DualPhi[R], DualPhi[L], h0, hp, hm);
}

// remap/UtilesDualRemap.cc:10-233 (A1), :401-644 (A2), :648-818 (PB), with
// remap/FacesLeftAndRight.h / FacesTopAndBottom.h
void ref_ctx::massFluxes(int sweep, bool horizontal, long p) {
  const int m = o.nbmat;
  const auto& FF = (sweep == 1) ? FluxFace1 : FluxFace2;
  const auto& DPF = (sweep == 1) ? DualphiFlux1 : DualphiFlux2;
  const auto& tab = horizontal ? VerticalFaceOfNode[p] : HorizontalFaceOfNode[p];
  // per face: the "plus" cell (right / above) and the "minus" cell
  // (left / below); slots of the face and of the far face inside each
  long cp[2], cm[2];
  int nf = 0;
  for (int s = 0; s < 2; ++s) {
    if (tab[s] == -1) continue;
    long b, fr;
    mesh.cellsOfFace(tab[s], b, fr);
    if (horizontal) cp[nf] = fr, cm[nf] = b;  // front = right, back = left
    else cp[nf] = b, cm[nf] = fr;             // back = above, front = below
    ++nf;
  }
  const int nbfaces = 2 * nf;
  // slots: horizontal: plus cell uses (left=1, right=2), minus cell (right=2, left=1)
  //        vertical:   plus cell uses (bottom=0, top=3), minus cell (top=3, bottom=0)
  const int pa = horizontal ? 1 : 0, pb = horizontal ? 2 : 3;
  const int ma = horizontal ? 2 : 3, mb = horizontal ? 1 : 0;
  const int comp = horizontal ? 0 : 1;
  double plus = 0., minus = 0.;
  for (int k = 0; k < m; ++k) {
    double pk = 0, mk = 0;
    if (o.methode_flux_masse == 0 || o.methode_flux_masse == 1) {
      const int e = (o.methode_flux_masse == 0) ? m + k : k;
      for (int s = 0; s < nf; ++s) {
        double tp = FF[cp[s]][pa][e] * outerFaceNormal[cp[s]][pa][comp] + FF[cp[s]][pb][e] * outerFaceNormal[cp[s]][pb][comp];
        double tm = FF[cm[s]][ma][e] * outerFaceNormal[cm[s]][ma][comp] + FF[cm[s]][mb][e] * outerFaceNormal[cm[s]][mb][comp];
        if (o.methode_flux_masse == 1) tp = rLagrange[cp[s]] * tp, tm = rLagrange[cm[s]] * tm;
        pk += tp;
        mk += tm;
      }
      if (nbfaces != 0) pk /= nbfaces, mk /= nbfaces;
    } else if (o.methode_flux_masse == 2) {
      int nbcell = 0;
      for (int s = 0; s < nf; ++s) {
        pk += DPF[cp[s]][m + k];
        mk += DPF[cm[s]][m + k];
        nbcell++;
      }
      if (nbcell != 0) pk /= nbcell, mk /= nbcell;
    }
    plus += pk;
    minus += mk;
  }
  if (horizontal) RightFluxMasse[p] = plus, LeftFluxMasse[p] = minus;
  else TopFluxMasse[p] = plus, BottomFluxMasse[p] = minus;
}

// remap/UtilesDualRemap.cc:234-397; side: 0 left, 1 right, 2 bottom, 3 top.
// Edge nodes (nb == -1) read the neighbour as 0: their mass flux is 0, so
// the product with the velocity vanishes as in the reference.
V2 ref_ctx::upwindVelocity(int side, long nb, long p, const VE& gradNb, const VE& grad0) const {
  const int order2 = o.projection_order - 1;
  const int a = (side <= 1) ? 0 : 1;
  const bool minusSide = (side == 0 || side == 2);
  V2 dpNb = {0.0, 0.0};
  double xNb = 0.0;
  if (nb >= 0) dpNb = {DualPhi[nb][0], DualPhi[nb][1]}, xNb = XLagrange[nb][a];
  const double u = minusSide ? 0.5 * (dpNb[a] + DualPhi[p][a]) : 0.5 * (DualPhi[p][a] + dpNb[a]);
  const int sign = (u > 0) ? 1 : -1;
  const double h = minusSide ? (XLagrange[p][a] - xNb) : (xNb - XLagrange[p][a]);
  double flux;
  if (side == 0) flux = LeftFluxMasse[p];
  else if (side == 1) flux = RightFluxMasse[p];
  else if (side == 2) flux = BottomFluxMasse[p];
  else flux = TopFluxMasse[p];
  const bool fromNb = minusSide ? (flux > 0) : (flux < 0);
  V2 r;
  for (int d = 0; d < 2; ++d) {
    const double base = fromNb ? dpNb[d] : DualPhi[p][d];
    const double g = fromNb ? gradNb[d] : grad0[d];
    r[d] = base + order2 * (0.5 * g * (sign * h - deltat_n * u));
  }
  return r;
}

// remap/DualPhaseRemap1.cc:18-136, remap/DualPhaseRemap2.cc:18-127
void ref_ctx::computeDualUremap(int sweep) {
  const bool xdir = (sweep == 1) ? x_then_y_n : !x_then_y_n;
  auto& gDP = (sweep == 1) ? gradDualPhi1 : gradDualPhi2;
  const auto& Uin = (sweep == 1) ? UDualLagrange : UDualremap1;
  auto& Uout = (sweep == 1) ? UDualremap1 : UDualremap2;
  const long N = mesh.nbNodes;
#pragma omp parallel for
  for (long p = 0; p < N; ++p) gDP[p] = (o.projection_order > 1) ? dualGradPhi(xdir, p) : ezero();
#pragma omp parallel for
  for (long p = 0; p < N; ++p) {
    massFluxes(sweep, xdir, p);
    const VE zero = ezero();
    if (xdir) {
      long L = mesh.leftNode(p), R = mesh.rightNode(p);
      Uout[p][2] = Uin[p][2] + LeftFluxMasse[p] - RightFluxMasse[p];
      LeftupwindVelocity[p] = upwindVelocity(0, L, p, L >= 0 ? gDP[L] : zero, gDP[p]);
      RightupwindVelocity[p] = upwindVelocity(1, R, p, R >= 0 ? gDP[R] : zero, gDP[p]);
      for (int d = 0; d < 2; ++d)
        Uout[p][d] = Uin[p][d] + LeftFluxMasse[p] * LeftupwindVelocity[p][d] -
                     RightFluxMasse[p] * RightupwindVelocity[p][d];
    } else {
      long B = mesh.bottomNode(p), T = mesh.topNode(p);
      Uout[p][2] = Uin[p][2] + BottomFluxMasse[p] - TopFluxMasse[p];
      BottomupwindVelocity[p] = upwindVelocity(2, B, p, B >= 0 ? gDP[B] : zero, gDP[p]);
      TopupwindVelocity[p] = upwindVelocity(3, T, p, T >= 0 ? gDP[T] : zero, gDP[p]);
      for (int d = 0; d < 2; ++d)
        Uout[p][d] = Uin[p][d] + BottomFluxMasse[p] * BottomupwindVelocity[p][d] -
                     TopFluxMasse[p] * TopupwindVelocity[p][d];
    }
  }
  if (sweep == 1) {
#pragma omp parallel for
    for (long p = 0; p < N; ++p) {
      DualPhi[p][0] = UDualremap1[p][0] / UDualremap1[p][2];
      DualPhi[p][1] = UDualremap1[p][1] / UDualremap1[p][2];
      DualPhi[p][2] = UDualremap1[p][2];
      DualPhi[p][3] = UDualremap1[p][3] / UDualremap1[p][2];
    }
  }
}

// vnr_remap/PhaseRemapFinal.cc:18-234
int ref_ctx::remapVariables() {
  x_then_y_nplus1 = !x_then_y_n;
  const int m = o.nbmat;
  const double thr = o.threshold;
  int bad = 0;
#pragma omp parallel for reduction(| : bad)
  for (long c = 0; c < mesh.nbCells; ++c) {
    const VE& U = Uremap2[c];
    const double vol = euler_volume[c];
    double volt = 0., masset = 0.;
    VM vn = {0.0, 0.0, 0.0};
    for (int k = 0; k < m; ++k) {
      vn[k] = U[k];
      volt += vn[k];
      masset += U[m + k];
    }
    double vtn = 0.;
    for (int k = 0; k < m; ++k) {
      vn[k] *= vol / volt;
      vtn += vn[k];
    }
    double sf = 0.;
    for (int k = 0; k < m; ++k) {
      fracvol_env[c][k] = vn[k] / vtn;
      if (fracvol_env[c][k] < thr) fracvol_env[c][k] = 0.;
      sf += fracvol_env[c][k];
    }
    for (int k = 0; k < m; ++k) fracvol_env[c][k] = fracvol_env[c][k] / sf;
    int matcell = 0, imatpure = -1;
    for (int k = 0; k < m; ++k)
      if (fracvol_env[c][k] > 0.) matcell++, imatpure = k;
    if (matcell > 1) mixte[c] = 1, pure[c] = -1;
    else mixte[c] = 0, pure[c] = imatpure;
    for (int k = 0; k < m; ++k) mass_fraction_env[c][k] = U[m + k] / masset;
    double fm = 0.;
    for (int k = 0; k < m; ++k) {
      if (fracvol_env[c][k] < thr) mass_fraction_env[c][k] = 0.;
      fm += mass_fraction_env[c][k];
    }
    for (int k = 0; k < m; ++k) mass_fraction_env[c][k] /= fm;
    VM dens = {0.0, 0.0, 0.0};
    double dn = 0.;
    for (int k = 0; k < m; ++k) {
      if (fracvol_env[c][k] > thr) dens[k] = U[m + k] / vn[k];
      dn += fracvol_env[c][k] * dens[k];
    }
    VM pe = {0.0, 0.0, 0.0};
    for (int k = 0; k < m; ++k)
      if ((fracvol_env[c][k] > thr) && (U[m + k] != 0.)) pe[k] = U[2 * m + k] / U[m + k];
    density_nplus1[c] = dn;
    energy_nplus1[c] = 0.;
    cell_mass[c] = euler_volume[c] * density_nplus1[c];
    for (int k = 0; k < m; ++k) cell_mass_env[c][k] = mass_fraction_env[c][k] * cell_mass[c];
    pseudo_nplus1[c] = U[3 * m + 3] / vol;
    for (int k = 0; k < m; ++k) {
      density_env_nplus1[c][k] = dens[k];
      energy_env_nplus1[c][k] = pe[k];
      energy_nplus1[c] += mass_fraction_env[c][k] * energy_env_nplus1[c][k];
    }
    for (int k = 0; k < m; ++k) {
      if (pe[k] < 0. || dens[k] < 0.) {
        density_env_nplus1[c][k] = 0.;
        energy_env_nplus1[c][k] = 0.;
        mass_fraction_env[c][k] = 0.;
        fracvol_env[c][k] = 0.;
      }
    }
    if (energy_nplus1[c] != energy_nplus1[c] || density_nplus1[c] != density_nplus1[c]) bad = 1;
    fracvol_env1[c] = fracvol_env[c][0];
    fracvol_env2[c] = fracvol_env[c][1];
    fracvol_env3[c] = fracvol_env[c][2];
  }
#pragma omp parallel for
  for (long p = 0; p < mesh.nbNodes; ++p) {
    double M = UDualremap2[p][2];
    node_velocity_nplus1[p][0] = UDualremap2[p][0] / M;
    node_velocity_nplus1[p][1] = UDualremap2[p][1] / M;
    x_velocity[p] = node_velocity_nplus1[p][0];
    y_velocity[p] = node_velocity_nplus1[p][1];
  }
  return bad ? MHY_ERR_NAN : MHY_OK;
}

// ---------------------------------------------------------------- time loop

void ref_ctx::swapAll() {
  std::swap(x_then_y_nplus1, x_then_y_n);
  std::swap(t_nplus1, t_n);
  std::swap(deltat_nplus1, deltat_n);
  std::swap(density_nplus1, density_n);
  std::swap(density_env_nplus1, density_env_n);
  std::swap(pressure_nplus1, pressure_n);
  std::swap(pressure_env_nplus1, pressure_env_n);
  std::swap(pseudo_nplus1, pseudo_n);
  std::swap(pseudo_env_nplus1, pseudo_env_n);
  std::swap(tau_nplus1, tau_n);
  std::swap(tau_env_nplus1, tau_env_n);
  std::swap(divu_nplus1, divu_n);
  std::swap(speed_nplus1, speed_n);
  std::swap(speed_env_nplus1, speed_env_n);
  std::swap(energy_nplus1, energy_n);
  std::swap(energy_env_nplus1, energy_env_n);
  std::swap(node_velocity_nplus1, node_velocity_n);
  if (o.avec_projection == 0) {
    std::swap(cell_coord_nplus1, cell_coord_n);
    std::swap(node_cellvolume_nplus1, node_cellvolume_n);
    std::swap(node_coord_nplus1, node_coord_n);
  }
}

// vnr_remap/VnrRemap.cc:130-222 (one pass of the do-while body)
int ref_ctx::iterate(bool dump, bool* continueLoop) {
  n++;
  if (o.sans_lagrange == 0) {
    computeCornerNormal();
    computeDeltaT();
  } else {
    deltat_nplus1 = deltat_n;
  }
  t_nplus1 = t_n + deltat_nplus1;
  if (dump) dumpVariables();
  if (o.sans_lagrange == 0) {
    updateVelocity();
    updateVelocityBoundaryConditions();
  } else {
    std::swap(node_velocity_n, node_velocity_nplus1);
  }
  updatePosition();
  updateCellPos();
  computeSubVol();
  updateRho();
  updatePeriodicBoundaryConditions();
  if (o.sans_lagrange == 0) {
    computeTau();
    computeDivU();
    computeArtificialViscosity();
    int st = updateEnergy();
    if (st != MHY_OK) return st;
    computeEOS();
    computePressionMoyenne();
  } else {
    std::swap(energy_n, energy_nplus1);
  }
  if (o.avec_projection == 1) {
    computeVariablesForRemap();
    computeCellQuantitesForRemap();
    computeFaceQuantitesForRemap();
    for (int s = 1; s <= 2; ++s) {
      computeGradPhiFace(s);
      computeGradPhi(s);
      computeUpwindFaceQuantitiesForProjection(s);
      computeUremap(s);
      computeDualUremap(s);
    }
    int st = remapVariables();
    if (st != MHY_OK) return st;
    computeNodeMass();
    computeEOS();
    computePressionMoyenne();
  }
  *continueLoop = (n + 1 < o.max_time_iterations && t_nplus1 < o.final_time);
  return MHY_OK;
}

// cppnabla/mesh/PvdFileWriter2D.cc:26-114: one VTP file of module
// <module> and its PVD collection (the point / cell data blocks only when
// they hold fields)
void ref_ctx::writeVtp(const std::string& module, long N, const V2* nodes, long C,
                       const std::map<std::string, const double*>& cellVars,
                       const std::map<std::string, const double*>& nodeVars,
                       std::map<double, std::string>& byTimes) {
  const std::string fileName = module + "." + std::to_string(nbCalls) + ".vtp";
  std::ofstream w(outdir + "/" + fileName);
  w << "<?xml version=\"1.0\"?>" << std::endl;
  w << "<VTKFile type=\"PolyData\">" << std::endl;
  w << "\t<PolyData>" << std::endl;
  w << "\t\t<Piece NumberOfPoints=\"" << N << "\" NumberOfPolys=\"" << C << "\">" << std::endl;
  w << "\t\t\t<Points>" << std::endl;
  w << "\t\t\t\t<DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"ascii\">" << std::endl;
  for (long r = 0; r < N; ++r) w << " " << nodes[r][0] << " " << nodes[r][1] << " " << 0.0;
  w << std::endl;
  w << "\t\t\t\t</DataArray>" << std::endl;
  w << "\t\t\t</Points>" << std::endl;
  w << "\t\t\t<Polys>" << std::endl;
  w << "\t\t\t\t<DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">" << std::endl;
  for (long c = 0; c < C; ++c) {
    w << " ";
    for (long id : mesh.nodesOfCell(c)) w << " " << id;
  }
  w << std::endl;
  w << "\t\t\t\t</DataArray>" << std::endl;
  w << "\t\t\t\t<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">" << std::endl;
  for (long c = 1; c <= C; ++c) w << " " << c * 4;
  w << std::endl;
  w << "\t\t\t\t</DataArray>" << std::endl;
  w << "\t\t\t</Polys>" << std::endl;
  auto block = [&](const char* tag, const std::map<std::string, const double*>& vars, long n) {
    if (vars.empty()) return;
    w << "\t\t\t<" << tag << " Scalars=\"" << vars.begin()->first << "\">" << std::endl;
    for (auto& kv : vars) {
      w << "\t\t\t\t<DataArray Name=\"" << kv.first << "\" type=\"Float32\" format=\"ascii\">" << std::endl;
      for (long r = 0; r < n; ++r) w << " " << kv.second[r];
      w << std::endl;
      w << "\t\t\t\t</DataArray>" << std::endl;
    }
    w << "\t\t\t</" << tag << ">" << std::endl;
  };
  block("PointData", nodeVars, N);
  block("CellData", cellVars, C);
  w << "\t\t</Piece>" << std::endl;
  w << "\t</PolyData>" << std::endl;
  w << "</VTKFile>" << std::endl;
  w.close();
  byTimes[t_n] = fileName;
  std::ofstream pvd(outdir + "/" + module + ".pvd");
  pvd << "<?xml version=\"1.0\"?>" << std::endl;
  pvd << "<VTKFile type=\"Collection\" version=\"0.1\">" << std::endl;
  pvd << "\t<Collection>" << std::endl;
  for (auto& kv : byTimes)
    pvd << "\t\t\t<DataSet timestep=\"" << kv.first << "\" group=\"\" part=\"0\" file=\"" << kv.second << "\"/>"
        << std::endl;
  pvd << "\t</Collection>" << std::endl;
  pvd << "</VTKFile>" << std::endl;
}

// Vnr::dumpVariables (vnr_remap/VnrRemap.cc:255-287)
void ref_ctx::dumpVariables() {
  nbCalls++;
  if (writerDisabled || !(t_n >= lastDump + o.output_time || t_n == 0.)) return;
  std::map<std::string, const double*> nodeVars, cellVars;
  nodeVars["VitesseX"] = x_velocity.data();
  nodeVars["VitesseY"] = y_velocity.data();
  cellVars["Pressure"] = pressure_n.data();
  cellVars["Density"] = density_n.data();
  cellVars["Energy"] = energy_n.data();
  if (o.nbmat > 1 && o.avec_projection == 1) cellVars["fracvol1"] = fracvol_env1.data();
  if (o.nbmat > 2 && o.avec_projection == 1) cellVars["fracvol2"] = fracvol_env2.data();
  writeVtp("VnrRemap", mesh.nbNodes, node_coord_n.data(), mesh.nbCells, cellVars, nodeVars, fileNameByTimes);
  lastDump = t_n;
}

// ================================================================ Eucclhyd
//
// The cell-centred Eucclhyd Lagrange scheme (eucclhyd_remap/), which runs
// the same cell sweeps as VNR (not the dual remap) on a remap vector that
// carries the cell momentum in components 3m, 3m+1 (and the kinetic energy
// in 3m+2 when the projection is conservative).

namespace {
typedef std::array<double, 4> M22;
// types/MathFunctions.h:41-47 (each row a dot product from 0.0)
inline V2 matVec(const M22& a, const V2& b) { return {dot2({a[0], a[1]}, b), dot2({a[2], a[3]}, b)}; }
// eucclhyd_remap/Utiles.cc:6-10 (types/MathFunctions.h:56-59 for det)
inline M22 inverse2(const M22& a) {
  const double alpha = 1.0 / (a[0] * a[3] - a[1] * a[2]);
  return {a[3] * alpha, -a[1] * alpha, -a[2] * alpha, a[0] * alpha};
}
// eucclhyd_remap/Utiles-Impl.h:4-12
inline M22 tens(const V2& a, const V2& b) { return {a[0] * b[0], a[0] * b[1], a[1] * b[0], a[1] * b[1]}; }
inline M22 madd(const M22& a, const M22& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2], a[3] + b[3]}; }
inline M22 mscale(double s, const M22& a) { return {s * a[0], s * a[1], s * a[2], s * a[3]}; }
constexpr double kEucCfl = 0.45;  // includes/GestionTemps.h:11 (never read from input)
}  // namespace

void ref_ctx::eucAllocate() {
  const size_t C = mesh.nbCells, N = mesh.nbNodes, F = mesh.nbFaces;
  zfill(e_lpc, C), zfill(e_nplus, C), zfill(e_nminus, C), zfill(e_lplus, C), zfill(e_lminus, C);
  zfill(e_perimeter, C), zfill(e_cell_deltat, C), zfill(x_cell_velocity, C), zfill(y_cell_velocity, C);
  zfill(cell_velocity_n, C), zfill(cell_velocity_nplus1, C), zfill(cell_velocity_n0, C);
  zfill(e_velgrad, C), zfill(e_pgrad_env, C), zfill(e_p_extrap, C), zfill(e_pe_extrap, C), zfill(e_v_extrap, C);
  zfill(e_F_n, C), zfill(e_F_nplus1, C), zfill(e_Fenv_n, C), zfill(e_Fenv_nplus1, C);
  zfill(e_M, C), zfill(e_Menv, C), zfill(e_nodeM, N), zfill(e_nodeG, N), zfill(XfLagrange, F);
  zfill(lagrange_volume, C);
}

// eucclhyd_remap/Init.cc:15-100 (a case without a rule keeps the
// includes/ConditionsLimites.h:22-35 defaults, code 0: zero velocity)
void ref_ctx::eucInitBoundaryConditions() {
  const int sym = MHY_BC_SYMMETRY, imp = MHY_BC_IMPOSED_VELOCITY;
  const V2 ex = {1.0, 0.0}, ey = {0.0, 1.0}, zero = {0.0, 0.0};
  const int t = o.test_case;
  if (t == MHY_CASE_SOD_X || t == MHY_CASE_SOD_Y || t == MHY_CASE_BI_SOD_X || t == MHY_CASE_BI_SOD_Y ||
      t == MHY_CASE_TRIPLE_POINT || t == MHY_CASE_BI_TRIPLE_POINT) {
    leftBC = sym, leftBCValue = ey, rightBC = sym, rightBCValue = ey;
    topBC = sym, topBCValue = ex, bottomBC = sym, bottomBCValue = ex;
  } else if (t == MHY_CASE_BI_SHOCK_BUBBLE) {
    leftBC = sym, leftBCValue = ey;
    rightBC = imp, rightBCValue = {-124.824, 0.0};
    rightFluxBC = 1;
    rightFluxBCValue = {1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, -124.824, 0.0, 250000, 0.0};
    topBC = sym, topBCValue = ex, bottomBC = sym, bottomBCValue = ex;
  } else if (t == MHY_CASE_SEDOV || t == MHY_CASE_BI_SEDOV || t == MHY_CASE_NOH || t == MHY_CASE_BI_NOH) {
    leftBC = sym, leftBCValue = ey;
    rightBC = imp, rightBCValue = zero;
    topBC = imp, topBCValue = zero;
    bottomBC = sym, bottomBCValue = ex;
  }
  FluxBC = leftFluxBC + rightFluxBC + bottomFluxBC + topFluxBC;
}

// computeCornerNormal, eucclhyd_remap/PhaseLagrange.cc:27-66 (on the
// set-up mesh; the time loop never forms them again)
void ref_ctx::eucCornerNormals() {
  for (long c = 0; c < mesh.nbCells; ++c) {
    auto nd = mesh.nodesOfCell(c);
    for (int s = 0; s < 4; ++s) {
      const V2& xp = node_coord_n[nd[s]];
      const V2& xpPlus = node_coord_n[nd[(s + 1) % 4]];
      const V2& xpMinus = node_coord_n[nd[(s + 3) % 4]];
      V2 np = {0.5 * (xpPlus[1] - xp[1]), 0.5 * (xp[0] - xpPlus[0])};
      const double lp = norm2(np);
      np = vdiv(np, lp);
      V2 nm = {0.5 * (xp[1] - xpMinus[1]), 0.5 * (xpMinus[0] - xp[0])};
      const double lm = norm2(nm);
      nm = vdiv(nm, lm);
      e_nplus[c][s] = np, e_lplus[c][s] = lp;
      e_nminus[c][s] = nm, e_lminus[c][s] = lm;
      e_lpc[c][s] = vadd(vmul(np, lp), vmul(nm, lm));
    }
  }
}

// eucclhyd_remap/Init.cc:106-167: polygon volume, centroid and perimeter
void ref_ctx::eucInitMeshGeometryForCells() {
  for (long c = 0; c < mesh.nbCells; ++c) {
    auto nd = mesh.nodesOfCell(c);
    double r11 = 0.0;
    for (int s = 0; s < 4; ++s) r11 = r11 + cross2(node_coord_n[nd[s]], node_coord_n[nd[(s + 1) % 4]]);
    const double vol = 0.5 * r11;
    V2 r12 = {0.0, 0.0};
    for (int s = 0; s < 4; ++s) {
      const V2& a = node_coord_n[nd[s]];
      const V2& b = node_coord_n[nd[(s + 1) % 4]];
      r12 = vadd(r12, vmul(vadd(a, b), cross2(a, b)));
    }
    cell_coord_n[c] = vmul(r12, 1.0 / (6.0 * vol));
    euler_volume[c] = vol;
    double r13 = 0.0;
    for (int s = 0; s < 4; ++s) r13 = r13 + norm2(vsub(node_coord_n[nd[s]], node_coord_n[nd[(s + 1) % 4]]));
    e_perimeter[c] = r13;
  }
}

// initCellInternalEnergy, initCellVelocity, initDensity
// (eucclhyd_remap/Init.cc:194-671).  Cases whose reference initialisation
// leaves the density unset (Noh, BiNoh, BiSedov) are refused.
int ref_ctx::eucInitState() {
  const long C = mesh.nbCells;
  const int t = o.test_case;
  switch (t) {
    case MHY_CASE_UNIT_TEST: case MHY_CASE_BI_UNIT_TEST: case MHY_CASE_SOD_X: case MHY_CASE_SOD_Y:
    case MHY_CASE_BI_SOD_X: case MHY_CASE_BI_SOD_Y: case MHY_CASE_BI_SHOCK_BUBBLE: case MHY_CASE_SEDOV:
    case MHY_CASE_TRIPLE_POINT: case MHY_CASE_BI_TRIPLE_POINT:
      break;
    default:
      return MHY_ERR_UNKNOWN_CASE;
  }
  const double* g = o.gammap;
  const V2 Xb = {0.320, 0.04};
  const double rb = 0.025;
  auto bubbleR = [&](const V2& x) { return std::sqrt((x[0] - Xb[0]) * (x[0] - Xb[0]) + (x[1] - Xb[1]) * (x[1] - Xb[1])); };
  for (long c = 0; c < C; ++c) {
    const V2& xc = cell_coord_n[c];
    VM& ee = energy_env_n0[c];
    ee = {0.0, 0.0, 0.0};
    double& e = energy_n0[c];
    // internal energy (Init.cc:194-385)
    if (t == MHY_CASE_UNIT_TEST || t == MHY_CASE_BI_UNIT_TEST) {
      const double p0 = 1., rho0 = 1., gamma = 1.4;
      const double e0 = p0 / ((gamma - 1.0) * rho0);
      e = e0;
      for (int k = 0; k < NMAT; ++k) ee[k] = e0;
    } else if (t == MHY_CASE_SOD_X || t == MHY_CASE_SOD_Y) {
      const double r = (t == MHY_CASE_SOD_X) ? xc[0] : xc[1];
      double pI, rI;
      if (r <= 0.5) pI = 1.0, rI = 1.0;
      else pI = 0.1, rI = 0.125;
      const double eI = pI / ((g[0] - 1.0) * rI);
      e = eI, ee[0] = eI;
    } else if (t == MHY_CASE_BI_SOD_X || t == MHY_CASE_BI_SOD_Y) {
      const double r = (t == MHY_CASE_BI_SOD_X) ? xc[0] : xc[1];
      if (r <= 0.5) {
        const double eI = 1.0 / ((g[0] - 1.0) * 1.0);
        ee[0] = eI, ee[1] = 0., e = eI;
      } else {
        const double eI = 0.1 / ((g[1] - 1.0) * 0.125);
        ee[0] = 0., ee[1] = eI, e = eI;
      }
    } else if (t == MHY_CASE_BI_SHOCK_BUBBLE) {
      double eI = 1.e5 / ((g[0] - 1.0) * 1.0);
      ee[0] = eI, ee[1] = 0., e = eI;
      if (bubbleR(xc) <= rb) {
        eI = 1.e5 / ((g[0] - 1.0) * 0.182);
        ee[0] = 0., ee[1] = eI, e = eI;
      }
    } else if (t == MHY_CASE_SEDOV) {
      bool center = false;
      const double pI = 1.e-6, rI = 1.;
      const double rmin = o.threshold;
      const double e1 = pI / ((g[0] - 1.0) * rI);
      for (long p : mesh.nodesOfCell(c))
        if (norm2(node_coord_n[p]) < rmin) center = true;
      if (center) e = e1 + 0.244816 / (o.dx * o.dy);
      else e = e1;
      ee[0] = e;
    } else if (t == MHY_CASE_TRIPLE_POINT) {
      double pI, rI;
      if (xc[0] <= 0.01) pI = 1.0, rI = 1.0;
      else if (xc[1] <= 0.015) pI = 0.1, rI = 1.;
      else pI = 0.1, rI = 0.1;
      const double eI = pI / ((g[0] - 1.0) * rI);
      e = eI, ee[0] = eI;
    } else {  // BiTriplePoint
      if (xc[0] <= 0.01) {
        const double eI = 1.0 / ((g[0] - 1.0) * 1.0);
        e = eI, ee[0] = eI;
      } else if (xc[1] <= 0.015) {
        const double eI = 0.1 / ((g[1] - 1.0) * 1.);
        e = eI, ee[1] = eI;
      } else {
        const double eI = 0.1 / ((g[2] - 1.0) * 0.1);
        e = eI, ee[2] = eI;
      }
    }
    // cell velocity (Init.cc:391-423); zero unless set
    V2& v = cell_velocity_n0[c];
    v = {0.0, 0.0};
    if (t == MHY_CASE_BI_SHOCK_BUBBLE && xc[0] >= 0.60) v = {-124.824, 0.0};
    if (t == MHY_CASE_UNIT_TEST || t == MHY_CASE_BI_UNIT_TEST) v = {1.0, 0.0};
    // density, fractions (Init.cc:429-671)
    VM& f = fracvol_env[c];
    VM& mf = mass_fraction_env[c];
    VM& re = density_env_n0[c];
    double& rho = density_n0[c];
    f = {0.0, 0.0, 0.0}, mf = {0.0, 0.0, 0.0}, re = {0.0, 0.0, 0.0};
    if (t == MHY_CASE_UNIT_TEST || (t == MHY_CASE_BI_UNIT_TEST && xc[0] <= 0.5)) {
      f[0] = 1., mf[0] = 1., rho = 1.0, re[0] = 1.0;
    } else if (t == MHY_CASE_BI_UNIT_TEST) {
      f[1] = 1., mf[1] = 1., rho = 1., re[1] = 1.0;
    } else if (t == MHY_CASE_SOD_X || t == MHY_CASE_SOD_Y) {
      const double r = (t == MHY_CASE_SOD_X) ? xc[0] : xc[1];
      f[0] = 1., mf[0] = 1.;
      const double d = (r <= 0.5) ? 1.0 : 0.125;
      rho = d, re[0] = d, re[1] = d;
    } else if (t == MHY_CASE_BI_SOD_X || t == MHY_CASE_BI_SOD_Y) {
      const double r = (t == MHY_CASE_BI_SOD_X) ? xc[0] : xc[1];
      if (r <= 0.5) f[0] = 1., mf[0] = 1., rho = 1.0, re[0] = 1.0;
      else f[1] = 1., mf[1] = 1., rho = 0.125, re[1] = 0.125;
    } else if (t == MHY_CASE_BI_SHOCK_BUBBLE) {
      re[0] = 1.0, rho = 1.0, f[0] = 1., mf[0] = 1.;
      if (bubbleR(xc) <= rb) {
        re[0] = 0.0, re[1] = 0.182, rho = 0.182;
        f[0] = 0., f[1] = 1., mf[0] = 0., mf[1] = 1.;
      }
    } else if (t == MHY_CASE_TRIPLE_POINT) {
      f[0] = 1., mf[0] = 1.;  // the material density stays unset (zero), as in the reference
      rho = (xc[0] <= 0.01) ? 1.0 : ((xc[1] <= 0.015) ? 1.0 : 0.1);
    } else if (t == MHY_CASE_BI_TRIPLE_POINT) {
      if (xc[0] <= 0.01) rho = 1.0, f[0] = 1., mf[0] = 1., re[0] = 1.0;
      else if (xc[1] <= 0.015) rho = 1.0, f[1] = 1., mf[1] = 1., re[1] = 1.0;
      else rho = 0.1, f[2] = 1., mf[2] = 1., re[2] = 0.1;
    } else {  // Sedov
      rho = 1.0, re[0] = 1.0;
    }
    if (o.nbmat == 1) f = {1., 0., 0.}, mf = {1., 0., 0.};
    fracvol_env1[c] = f[0], fracvol_env2[c] = f[1], fracvol_env3[c] = f[2];
    x_cell_velocity[c] = v[0], y_cell_velocity[c] = v[1];
    int matcell = 0, imatpure = -1;
    for (int k = 0; k < NMAT; ++k)
      if (f[k] > o.threshold) matcell++, imatpure = k;
    if (matcell > 1) mixte[c] = 1, pure[c] = -1;
    else mixte[c] = 0, pure[c] = imatpure;
  }
  return MHY_OK;
}

// eucclhyd_remap/Init.cc:721-760 (t = 0 state and the case's initial time
// step; GestTemps is value-initialised, so a case without one starts at 0)
void ref_ctx::eucSetUpTimeLoopN() {
  node_velocity_n = node_velocity_n0;
  density_n = density_n0, density_env_n = density_env_n0;
  cell_velocity_n = cell_velocity_n0;
  energy_n = energy_n0, energy_env_n = energy_env_n0;
  const int t = o.test_case;
  deltat_n = 0.0;
  if (t == MHY_CASE_SOD_X || t == MHY_CASE_SOD_Y || t == MHY_CASE_BI_SOD_X || t == MHY_CASE_BI_SOD_Y ||
      t == MHY_CASE_SEDOV)
    deltat_n = 1.0e-4;
  else if (t == MHY_CASE_BI_SHOCK_BUBBLE)
    deltat_n = 1.0e-7;
  else if (t == MHY_CASE_TRIPLE_POINT || t == MHY_CASE_BI_TRIPLE_POINT)
    deltat_n = 1.0e-5;
  t_n = 0.0;
}

// Eucclhyd::simulate set-up (eucclhyd_remap/EucclhydRemap.cc:273-286)
int ref_ctx::eucPrologue() {
  for (int k = 0; k < o.nbmat; ++k)
    if (o.eos[k] != MHY_EOS_PERFECT_GAS && o.eos[k] != MHY_EOS_VOID) return MHY_ERR_ARG;
  if (o.space_order != 1 && o.space_order != 2) return MHY_ERR_ARG;
  node_coord_n = node_coord_n0;
  eucCornerNormals();
  eucInitMeshGeometryForCells();
  eucInitBoundaryConditions();
  int st = eucInitState();
  if (st != MHY_OK) return st;
  cell_coord_n0 = cell_coord_n;
  initMeshGeometryForFaces();  // same rule as vnr_remap/Init.cc:517-554 (Init.cc:677-715)
  eucSetUpTimeLoopN();
  return MHY_OK;
}

// computeEOS (PhaseLagrange.cc:70-105): perfect gas and void
void ref_ctx::eucComputeEOS() {
  for (int k = 0; k < o.nbmat; ++k) {
    const double g = o.gammap[k];
    for (long c = 0; c < mesh.nbCells; ++c) {
      if (o.eos[k] == MHY_EOS_PERFECT_GAS) {
        pressure_env_n[c][k] = (g - 1.0) * density_env_n[c][k] * energy_env_n[c][k];
        speed_env_n[c][k] = (density_env_n[c][k] > 0.) ? std::sqrt(g * pressure_env_n[c][k] / density_env_n[c][k]) : 1.e-20;
      } else {
        pressure_env_n[c][k] = 0.;
        speed_env_n[c][k] = 1.e-20;
      }
    }
  }
}

// computePressionMoyenne (PhaseLagrange.cc:139-154): the sound speed keeps
// its previous value in the max unless the density is positive
void ref_ctx::eucPressionMoyenne() {
  for (long c = 0; c < mesh.nbCells; ++c) {
    pressure_n[c] = 0.;
    for (int k = 0; k < o.nbmat; ++k) {
      pressure_n[c] += fracvol_env[c][k] * pressure_env_n[c][k];
      speed_n[c] = smax(speed_n[c], speed_env_n[c][k]);
    }
    if (density_n[c] > 0.) speed_n[c] = std::sqrt(o.gammap[0] * pressure_n[c] / density_n[c]);
  }
}

// computeGradients, second-order part (PhaseLagrange.cc:199-244): velocity
// gradient from the node velocities at n, material pressure gradients from
// the corner forces at n (the particle-only reductions are not formed)
void ref_ctx::eucGradients() {
  for (long c = 0; c < mesh.nbCells; ++c) {
    auto nd = mesh.nodesOfCell(c);
    M22 r = {0.0, 0.0, 0.0, 0.0};
    for (int s = 0; s < 4; ++s) r = madd(r, tens(node_velocity_n[nd[s]], e_lpc[c][s]));
    const double V = euler_volume[c];
    e_velgrad[c] = {r[0] / V, r[1] / V, r[2] / V, r[3] / V};
    for (int k = 0; k < NMAT; ++k) {
      V2 gk = {0.0, 0.0};
      for (int s = 0; s < 4; ++s) gk = vadd(gk, e_Fenv_n[c][s][k]);
      e_pgrad_env[c][k] = vdiv(gk, V);
    }
  }
}

// computeMass (PhaseLagrange.cc:251-260)
void ref_ctx::eucMass() {
  for (long c = 0; c < mesh.nbCells; ++c) {
    cell_mass[c] = density_n[c] * euler_volume[c];
    for (int k = 0; k < o.nbmat; ++k) cell_mass_env[c][k] = mass_fraction_env[c][k] * cell_mass[c];
  }
}

// computeDissipationMatrix (PhaseLagrange.cc:266-296)
void ref_ctx::eucDissipationMatrix() {
  for (long c = 0; c < mesh.nbCells; ++c)
    for (int s = 0; s < 4; ++s) {
      const M22 cm = madd(mscale(e_lplus[c][s], tens(e_nplus[c][s], e_nplus[c][s])),
                          mscale(e_lminus[c][s], tens(e_nminus[c][s], e_nminus[c][s])));
      e_M[c][s] = mscale(density_n[c] * speed_n[c], cm);
      for (int k = 0; k < NMAT; ++k) e_Menv[c][s][k] = mscale(density_env_n[c][k] * speed_n[c], cm);
    }
}

// computem_cell_deltat (PhaseLagrange.cc:302-316)
void ref_ctx::eucCellDeltat() {
  for (long c = 0; c < mesh.nbCells; ++c) {
    if (o.avec_projection == 1)
      e_cell_deltat[c] = euler_volume[c] / (e_perimeter[c] * (norm2(cell_velocity_n[c]) + speed_n[c]));
    else
      e_cell_deltat[c] = euler_volume[c] / (e_perimeter[c] * speed_n[c]);
  }
}

// extrapolateValue (PhaseLagrange.cc:322-478): second order extrapolates
// the material pressures and the velocity to each corner, clamped to the
// range of the node's cells (max reductions start at DBL_MIN, the smallest
// positive double, as the reference's numeric_limits<double>::min())
void ref_ctx::eucExtrapolate() {
  const int m = o.nbmat;
  if (o.space_order == 1) {
    for (long c = 0; c < mesh.nbCells; ++c)
      for (int s = 0; s < 4; ++s) {
        e_v_extrap[c][s] = cell_velocity_n[c];
        e_p_extrap[c][s] = pressure_n[c];
        for (int k = 0; k < m; ++k) e_pe_extrap[c][s][k] = pressure_env_n[c][k];
      }
    return;
  }
  const double big = std::numeric_limits<double>::max(), tiny = std::numeric_limits<double>::min();
  for (long p = 0; p < mesh.nbNodes; ++p) {
    long cells[4];
    int slot[4];
    const int nc = mesh.cellsOfNode(p, cells, slot);
    double minP[NMAT], maxP[NMAT];
    for (int k = 0; k < NMAT; ++k) minP[k] = big, maxP[k] = tiny;
    for (int q = 0; q < nc; ++q)
      for (int k = 0; k < NMAT; ++k) minP[k] = smin(minP[k], pressure_env_n[cells[q]][k]);
    for (int q = 0; q < nc; ++q)
      for (int k = 0; k < NMAT; ++k) maxP[k] = smax(maxP[k], pressure_env_n[cells[q]][k]);
    double minVx = big, maxVx = tiny, minVy = big, maxVy = tiny;
    for (int q = 0; q < nc; ++q) minVx = smin(minVx, cell_velocity_n[cells[q]][0]);
    for (int q = 0; q < nc; ++q) maxVx = smax(maxVx, cell_velocity_n[cells[q]][0]);
    for (int q = 0; q < nc; ++q) minVy = smin(minVy, cell_velocity_n[cells[q]][1]);
    for (int q = 0; q < nc; ++q) maxVy = smax(maxVy, cell_velocity_n[cells[q]][1]);
    for (int q = 0; q < nc; ++q) {
      const long c = cells[q];
      const int s = slot[q];
      const V2 d = vsub(node_coord_n[p], cell_coord_n[c]);
      for (int k = 0; k < NMAT; ++k) {
        const double ptmp = pressure_env_n[c][k] + dot2(e_pgrad_env[c][k], d);
        e_pe_extrap[c][s][k] = smax(smin(maxP[k], ptmp), minP[k]);
      }
      e_p_extrap[c][s] = 0.;
      for (int k = 0; k < m; ++k) e_p_extrap[c][s] += fracvol_env[c][k] * e_pe_extrap[c][s][k];
      const V2 vt = vadd(cell_velocity_n[c], matVec(e_velgrad[c], d));
      e_v_extrap[c][s] = {smax(smin(maxVx, vt[0]), minVx), smax(smin(maxVy, vt[1]), minVy)};
    }
  }
}

// computeG (PhaseLagrange.cc:484-504)
void ref_ctx::eucNodeG() {
  for (long p = 0; p < mesh.nbNodes; ++p) {
    long cells[4];
    int slot[4];
    const int nc = mesh.cellsOfNode(p, cells, slot);
    V2 r = {0.0, 0.0};
    for (int q = 0; q < nc; ++q) {
      const long c = cells[q];
      const int s = slot[q];
      r = vadd(r, vadd(matVec(e_M[c][s], e_v_extrap[c][s]), vmul(e_lpc[c][s], e_p_extrap[c][s])));
    }
    e_nodeG[p] = r;
  }
}

// computeNodeDissipationMatrixAndG (PhaseLagrange.cc:511-526)
void ref_ctx::eucNodeDissipation() {
  for (long p = 0; p < mesh.nbNodes; ++p) {
    long cells[4];
    int slot[4];
    const int nc = mesh.cellsOfNode(p, cells, slot);
    M22 r = {0.0, 0.0, 0.0, 0.0};
    for (int q = 0; q < nc; ++q) r = madd(r, e_M[cells[q]][slot[q]]);
    e_nodeM[p] = r;
  }
}

// computedeltat (EucclhydRemap.cc:203-227): Kokkos::Min keeps the smaller
// value (a NaN never replaces it)
int ref_ctx::eucDeltat() {
  double red = std::numeric_limits<double>::max();
  for (long c = 0; c < mesh.nbCells; ++c)
    if (e_cell_deltat[c] < red) red = e_cell_deltat[c];
  deltat_nplus1 = smin(kEucCfl * red, deltat_n * 1.05);
  if (deltat_nplus1 < o.deltat_min) return MHY_ERR_DT_MIN;
  return MHY_OK;
}

// nodeVelocityBoundaryCondition (ConditionsLimites.cc:109-123)
V2 ref_ctx::eucNodeBC(int bc, const V2& bcv, const M22& Mp, const V2& Gp) const {
  if (bc == MHY_BC_SYMMETRY) return vmul(bcv, dot2(Gp, bcv) / dot2(matVec(Mp, bcv), bcv));
  if (bc == MHY_BC_IMPOSED_VELOCITY) return bcv;
  if (bc == MHY_BC_FREE_SURFACE) return matVec(inverse2(Mp), Gp);
  return {0.0, 0.0};
}

// nodeVelocityBoundaryConditionCorner (ConditionsLimites.cc:125-145)
V2 ref_ctx::eucNodeBCCorner(int bc1, const V2& bcv1, int bc2, const V2& bcv2, const M22& Mp, const V2& Gp) const {
  if (bc1 == MHY_BC_SYMMETRY && bc2 == MHY_BC_SYMMETRY) {
    if (std::fabs(std::fabs(dot2(bcv1, bcv2)) - norm2(bcv1) * norm2(bcv2)) < 1.0E-8)
      return vmul(bcv1, dot2(Gp, bcv1) / dot2(matVec(Mp, bcv1), bcv1));
    return {0.0, 0.0};
  }
  if (bc1 == MHY_BC_IMPOSED_VELOCITY && bc2 == MHY_BC_IMPOSED_VELOCITY) return vmul(vadd(bcv1, bcv2), 0.5);
  if (bc1 == MHY_BC_FREE_SURFACE && bc2 == MHY_BC_FREE_SURFACE) return matVec(inverse2(Mp), Gp);
  return {0.0, 0.0};
}

// computeBoundaryNodeVelocities (ConditionsLimites.cc:19-107: sides left,
// right, top, bottom with their corners, then the four corners) and
// computeNodeVelocity (PhaseLagrange.cc:532-542) for the inner nodes
void ref_ctx::eucNodeVelocities() {
  const long nx = mesh.nx, ny = mesh.ny;
  auto& v = node_velocity_nplus1;
  auto at = [&](long j, long i) { return mesh.node(j, i); };
  for (long j = 0; j <= ny; ++j) v[at(j, 0)] = eucNodeBC(leftBC, leftBCValue, e_nodeM[at(j, 0)], e_nodeG[at(j, 0)]);
  for (long j = 0; j <= ny; ++j)
    v[at(j, nx)] = eucNodeBC(rightBC, rightBCValue, e_nodeM[at(j, nx)], e_nodeG[at(j, nx)]);
  for (long i = 0; i <= nx; ++i) v[at(ny, i)] = eucNodeBC(topBC, topBCValue, e_nodeM[at(ny, i)], e_nodeG[at(ny, i)]);
  for (long i = 0; i <= nx; ++i)
    v[at(0, i)] = eucNodeBC(bottomBC, bottomBCValue, e_nodeM[at(0, i)], e_nodeG[at(0, i)]);
  long p = at(ny, 0);
  v[p] = eucNodeBCCorner(topBC, topBCValue, leftBC, leftBCValue, e_nodeM[p], e_nodeG[p]);
  p = at(ny, nx);
  v[p] = eucNodeBCCorner(topBC, topBCValue, rightBC, rightBCValue, e_nodeM[p], e_nodeG[p]);
  p = at(0, 0);
  v[p] = eucNodeBCCorner(bottomBC, bottomBCValue, leftBC, leftBCValue, e_nodeM[p], e_nodeG[p]);
  p = at(0, nx);
  v[p] = eucNodeBCCorner(bottomBC, bottomBCValue, rightBC, rightBCValue, e_nodeM[p], e_nodeG[p]);
  for (long q = 0; q < mesh.nbNodes; ++q)
    if (mesh.innerNode(q)) v[q] = matVec(inverse2(e_nodeM[q]), e_nodeG[q]);
}

// computeFaceVelocity (PhaseLagrange.cc:548-566)
void ref_ctx::eucFaceVelocity() {
  for (long f = 0; f < mesh.nbFaces; ++f) {
    V2 r = {0.0, 0.0};
    r = vadd(r, node_velocity_nplus1[mesh.firstNode(f)]);
    r = vadd(r, node_velocity_nplus1[mesh.secondNode(f)]);
    faceNormalVelocity[f] = dot2(vmul(r, 0.5), faceNormal[f]);
  }
}

// computeLagrangePosition (PhaseLagrange.cc:573-618): the nodes move with
// the time step of step n; without remap the perimeter of the start-of-step
// mesh is kept for the next step's time step
void ref_ctx::eucLagrangePosition() {
  for (long p = 0; p < mesh.nbNodes; ++p) XLagrange[p] = vadd(node_coord_n[p], vmul(node_velocity_nplus1[p], deltat_n));
  for (long f = 0; f < mesh.nbFaces; ++f) {
    const long n1 = mesh.firstNode(f), n2 = mesh.secondNode(f);
    XfLagrange[f] = vmul(vadd(XLagrange[n1], XLagrange[n2]), 0.5);
    faceLengthLagrange[f] = norm2(vsub(XLagrange[n2], XLagrange[n1]));
  }
  if (o.avec_projection == 0) {
    for (long c = 0; c < mesh.nbCells; ++c) {
      auto nd = mesh.nodesOfCell(c);
      double r = 0.0;
      for (int s = 0; s < 4; ++s) r = r + norm2(vsub(node_coord_n[nd[s]], node_coord_n[nd[(s + 1) % 4]]));
      e_perimeter[c] = r;
    }
  }
}

// computeSubCellForce (PhaseLagrange.cc:625-666)
void ref_ctx::eucSubCellForce() {
  for (long c = 0; c < mesh.nbCells; ++c) {
    auto nd = mesh.nodesOfCell(c);
    for (int s = 0; s < 4; ++s) {
      const V2 dv = vsub(node_velocity_nplus1[nd[s]], e_v_extrap[c][s]);
      const V2& l = e_lpc[c][s];
      e_F_nplus1[c][s] = vadd(vmul(l, -e_p_extrap[c][s]), matVec(e_M[c][s], dv));
      for (int k = 0; k < NMAT; ++k)
        e_Fenv_nplus1[c][s][k] = vadd(vmul(l, -e_pe_extrap[c][s][k]), matVec(e_Menv[c][s][k], dv));
    }
  }
}

// computeLagrangeVolumeAndCenterOfGravity (PhaseLagrange.cc:671-715)
void ref_ctx::eucLagrangeGeometry() {
  for (long c = 0; c < mesh.nbCells; ++c) {
    auto nd = mesh.nodesOfCell(c);
    double r6 = 0.0;
    for (int s = 0; s < 4; ++s) r6 = r6 + cross2(XLagrange[nd[s]], XLagrange[nd[(s + 1) % 4]]);
    const double vol = 0.5 * r6;
    vLagrange[c] = vol;
    for (int k = 0; k < o.nbmat; ++k) lagrange_volume[c][k] = fracvol_env[c][k] * vol;
    V2 r7 = {0.0, 0.0};
    for (int s = 0; s < 4; ++s) {
      const V2& a = XLagrange[nd[s]];
      const V2& b = XLagrange[nd[(s + 1) % 4]];
      r7 = vadd(r7, vmul(vadd(a, b), cross2(a, b)));
    }
    XcLagrange[c] = vmul(r7, 1.0 / (6.0 * vol));
  }
}

// computeFacedeltaxLagrange (PhaseLagrange.cc:721-739)
void ref_ctx::eucFaceDeltax() {
  for (long f = 0; f < mesh.nbFaces; ++f) {
    if (!mesh.innerFace(f)) continue;
    long b, fr;
    mesh.cellsOfFace(f, b, fr);
    deltaxLagrange[f] = dot2(vsub(XcLagrange[fr], XcLagrange[b]), faceNormal[f]);
  }
}

// updateCellCenteredLagrangeVariables (PhaseLagrange.cc:746-1011)
void ref_ctx::eucCellUpdate() {
  const int m = o.nbmat;
  const double dt = deltat_n, thr = o.threshold;
  for (long c = 0; c < mesh.nbCells; ++c) {
    auto nd = mesh.nodesOfCell(c);
    double r2 = 0.0;
    for (int s = 0; s < 4; ++s) r2 = r2 + dot2(e_lpc[c][s], node_velocity_nplus1[nd[s]]);
    const double rhoL = 1 / (1 / density_n[c] + dt / cell_mass[c] * r2);
    V2 r3 = {0.0, 0.0};
    for (int s = 0; s < 4; ++s) r3 = vadd(r3, e_F_nplus1[c][s]);
    const V2 vc = cell_velocity_n[c];
    const V2 vL = vadd(vc, vdiv(vmul(r3, dt), cell_mass[c]));
    double r4 = 0.0;
    VM pr4 = {0.0, 0.0, 0.0};
    for (int s = 0; s < 4; ++s) {
      const V2 w = vsub(node_velocity_nplus1[nd[s]], vmul(vadd(vc, vL), 0.5));
      r4 = r4 + dot2(e_F_nplus1[c][s], w);
      for (int k = 0; k < NMAT; ++k) pr4[k] = pr4[k] + dot2(e_Fenv_nplus1[c][s][k], w);
    }
    (void)r4;  // eLagrange is not used by the reference
    VM peL = {0.0, 0.0, 0.0};
    const VM& f = fracvol_env[c];
    const VM& mf = mass_fraction_env[c];
    for (int k = 0; k < m; ++k)
      if (f[k] > thr && cell_mass_env[c][k] != 0.)
        peL[k] = energy_env_n[c][k] + f[k] * dt / cell_mass_env[c][k] * pr4[k];
    VE& U = ULagrange[c];
    const double vol = vLagrange[c];
    for (int k = 0; k < m; ++k) {
      U[k] = lagrange_volume[c][k];
      U[m + k] = mf[k] * vol * rhoL;
      U[2 * m + k] = mf[k] * vol * rhoL * peL[k];
    }
    U[3 * m] = vol * rhoL * vL[0];
    U[3 * m + 1] = vol * rhoL * vL[1];
    if (o.projection_conservative == 1) U[3 * m + 2] = 0.5 * vol * rhoL * (vL[0] * vL[0] + vL[1] * vL[1]);
    if (o.avec_projection == 0) {
      cell_velocity_nplus1[c] = vL;
      density_nplus1[c] = 0.;
      energy_nplus1[c] = 0.;
      for (int k = 0; k < m; ++k) {
        density_nplus1[c] += mf[k] * rhoL;
        density_env_nplus1[c][k] = (f[k] > thr) ? mf[k] * rhoL / f[k] : 0.;
        energy_nplus1[c] += mf[k] * peL[k];
        energy_env_nplus1[c][k] = (f[k] > thr) ? peL[k] : 0.;
      }
      fracvol_env1[c] = f[0], fracvol_env2[c] = f[1], fracvol_env3[c] = f[2];
      if (vL[0] > 0.) x_cell_velocity[c] = smax(vL[0], thr);
      if (vL[0] < 0.) x_cell_velocity[c] = smin(vL[0], -thr);
      if (vL[1] > 0.) y_cell_velocity[c] = smax(vL[1], thr);
      if (vL[1] < 0.) y_cell_velocity[c] = smin(vL[1], -thr);
    }
    VE& P = Phi[c];
    if (o.pp == 1) {
      double sv = 0.;
      for (int k = 0; k < m; ++k) sv += U[k];
      double sm = 0.;
      for (int k = 0; k < m; ++k) {
        P[k] = U[k] / sv;
        P[m + k] = (U[k] != 0.) ? U[m + k] / U[k] : 0.;
        sm += U[m + k];
      }
      P[3 * m] = U[3 * m] / sm;
      P[3 * m + 1] = U[3 * m + 1] / sm;
      for (int k = 0; k < m; ++k) P[2 * m + k] = (U[m + k] != 0.) ? U[2 * m + k] / U[m + k] : 0.;
      if (o.projection_conservative == 1) P[3 * m + 2] = U[3 * m + 2] / sm;
    } else {
      for (int e = 0; e < NEQ; ++e) P[e] = U[e] / vol;
    }
  }
}

// remapCellcenteredVariable (eucclhyd_remap/PhaseRemapFinal.cc:17-232)
int ref_ctx::eucRemapFinal() {
  x_then_y_nplus1 = !x_then_y_n;
  const int m = o.nbmat;
  const double thr = o.threshold;
  int bad = 0;
  for (long c = 0; c < mesh.nbCells; ++c) {
    const VE& U = Uremap2[c];
    const double vol = euler_volume[c];
    double volt = 0., masset = 0.;
    VM vn = {0.0, 0.0, 0.0};
    for (int k = 0; k < m; ++k) {
      vn[k] = U[k];
      volt += vn[k];
      masset += U[m + k];
    }
    double vtn = 0.;
    for (int k = 0; k < m; ++k) {
      vn[k] *= vol / volt;
      vtn += vn[k];
    }
    VM& f = fracvol_env[c];
    double sf = 0.;
    for (int k = 0; k < m; ++k) {
      f[k] = vn[k] / vtn;
      if (f[k] < thr) f[k] = 0.;
      sf += f[k];
    }
    for (int k = 0; k < m; ++k) f[k] = f[k] / sf;
    int matcell = 0, imatpure = -1;
    for (int k = 0; k < m; ++k)
      if (f[k] > 0.) matcell++, imatpure = k;
    if (matcell > 1) mixte[c] = 1, pure[c] = -1;
    else mixte[c] = 0, pure[c] = imatpure;
    VM& mf = mass_fraction_env[c];
    for (int k = 0; k < m; ++k) mf[k] = U[m + k] / masset;
    double fm = 0.;
    for (int k = 0; k < m; ++k) {
      if (f[k] < thr) mf[k] = 0.;
      fm += mf[k];
    }
    for (int k = 0; k < m; ++k) mf[k] /= fm;
    VM dens = {0.0, 0.0, 0.0};
    double dn = 0.;
    for (int k = 0; k < m; ++k) {
      if (f[k] > thr) dens[k] = U[m + k] / vn[k];
      dn += f[k] * dens[k];
    }
    const V2 v = {U[3 * m] / (dn * vol), U[3 * m + 1] / (dn * vol)};
    VM pe = {0.0, 0.0, 0.0};
    for (int k = 0; k < m; ++k)
      if ((f[k] > thr) && (U[m + k] != 0.)) pe[k] = U[2 * m + k] / U[m + k];
    density_nplus1[c] = dn;
    cell_velocity_nplus1[c] = v;
    energy_nplus1[c] = 0.;
    double dec = 0.;
    if (o.projection_conservative == 1) dec = U[3 * m + 2] / masset - 0.5 * (v[0] * v[0] + v[1] * v[1]);
    for (int k = 0; k < m; ++k) {
      density_env_nplus1[c][k] = dens[k];
      energy_env_nplus1[c][k] = pe[k];
      energy_env_nplus1[c][k] += dec;
      energy_nplus1[c] += mf[k] * energy_env_nplus1[c][k];
    }
    for (int k = 0; k < m; ++k) {
      if (pe[k] < 0. || dens[k] < 0.) {
        density_env_nplus1[c][k] = 0.;
        energy_env_nplus1[c][k] = 0.;
        mf[k] = 0.;
        f[k] = 0.;
      }
    }
    if (energy_nplus1[c] != energy_nplus1[c] || density_nplus1[c] != density_nplus1[c]) bad = 1;
    fracvol_env1[c] = f[0], fracvol_env2[c] = f[1], fracvol_env3[c] = f[2];
    if (v[0] > 0.) x_cell_velocity[c] = smax(v[0], thr);
    if (v[0] < 0.) x_cell_velocity[c] = smin(v[0], -thr);
    if (v[1] > 0.) y_cell_velocity[c] = smax(v[1], thr);
    if (v[1] < 0.) y_cell_velocity[c] = smin(v[1], -thr);
  }
  return bad ? MHY_ERR_NAN : MHY_OK;
}

// Eucclhyd::executeTimeLoopN body (EucclhydRemap.cc:84-142)
int ref_ctx::eucIterate(bool dump, bool* continueLoop) {
  n++;
  eucComputeEOS();
  eucPressionMoyenne();
  if (o.space_order == 2) eucGradients();
  eucMass();
  eucDissipationMatrix();
  eucCellDeltat();
  if (dump) eucDump();
  eucExtrapolate();
  eucNodeG();
  eucNodeDissipation();
  int st = eucDeltat();
  if (st != MHY_OK) return st;
  eucNodeVelocities();
  t_nplus1 = t_n + deltat_nplus1;
  eucFaceVelocity();
  eucLagrangePosition();
  eucSubCellForce();
  eucLagrangeGeometry();
  eucFaceDeltax();
  eucCellUpdate();
  if (o.avec_projection == 1) {
    for (int s = 1; s <= 2; ++s) {
      computeGradPhiFace(s);
      computeGradPhi(s);
      computeUpwindFaceQuantitiesForProjection(s);
      computeUremap(s);
    }
    st = eucRemapFinal();
    if (st != MHY_OK) return st;
  }
  *continueLoop = (n + 1 < o.max_time_iterations && t_nplus1 < o.final_time);
  return MHY_OK;
}

// the n / n+1 swaps of EucclhydRemap.cc:143-168
void ref_ctx::eucSwap() {
  std::swap(x_then_y_nplus1, x_then_y_n);
  std::swap(t_nplus1, t_n);
  std::swap(deltat_nplus1, deltat_n);
  std::swap(node_velocity_nplus1, node_velocity_n);
  std::swap(density_nplus1, density_n);
  std::swap(density_env_nplus1, density_env_n);
  std::swap(cell_velocity_nplus1, cell_velocity_n);
  std::swap(energy_nplus1, energy_n);
  std::swap(energy_env_nplus1, energy_env_n);
  std::swap(e_F_nplus1, e_F_n);
  std::swap(e_Fenv_nplus1, e_Fenv_n);
  if (o.avec_projection == 0) {
    std::swap(vLagrange, euler_volume);
    std::swap(XLagrange, node_coord_n);
    std::swap(XfLagrange, Xf);
    std::swap(faceLengthLagrange, faceLength);
    std::swap(XcLagrange, cell_coord_n);
  }
}

// dumpVariables (EucclhydRemap.cc:24-62): the cell fields, and the empty
// particle file of the disabled particle scheme
void ref_ctx::eucDump() {
  nbCalls++;
  if (writerDisabled || !(t_n >= lastDump + o.output_time || t_n == 0.)) return;
  std::map<std::string, const double*> cellVars, nodeVars, partVars;
  cellVars["Pressure"] = pressure_n.data();
  cellVars["Density"] = density_n.data();
  cellVars["F1"] = fracvol_env1.data();
  cellVars["F2"] = fracvol_env2.data();
  cellVars["F3"] = fracvol_env3.data();
  cellVars["VelocityX"] = x_cell_velocity.data();
  cellVars["VelocityY"] = y_cell_velocity.data();
  cellVars["Energy"] = energy_n.data();
  partVars["VolumePart"] = nullptr;
  partVars["VxPart"] = nullptr;
  partVars["VyPart"] = nullptr;
  writeVtp("EucclhydRemap", mesh.nbNodes, node_coord_n.data(), mesh.nbCells, cellVars, nodeVars, fileNameByTimes);
  writeVtp("Particules", 0, nullptr, 0, cellVars, partVars, particleFileNameByTimes);
  lastDump = t_n;
}

// ----------------------------------------------------------------- C API

extern "C" {

void ref_config_default(mhy_config* c) {
  std::memset(c, 0, sizeof(*c));
  c->test_case = -1;
  c->nx = 30, c->ny = 30;
  c->dx = 1.2 / 30, c->dy = 1.2 / 30;
  c->final_time = 1.0, c->output_time = 1.0;
  c->deltat_init = 0.0, c->deltat_min = 1.0E-10;
  c->max_time_iterations = 500000000;
  c->nbmat = -1;
  for (int k = 0; k < MHY_NBMATMAX; ++k) c->eos[k] = MHY_EOS_PERFECT_GAS, c->gammap[k] = 1.4;
  c->gamma = 1.4;
  c->schema = -1;
  c->space_order = 2;
  c->avec_projection = 1;
  c->projection_conservative = 0;
  c->threshold = 1.0E-16;
  c->projection_order = 2;
  c->limiter = -1, c->limiter_pure = -1;
  c->pp = 0, c->pp_mixte = 0, c->pp_complet = 0;
  c->pseudo_centree = 1;
  c->methode_flux_masse = 0;
  c->sans_lagrange = 0;
  c->abort_on_energy = 1;
}

// lecture_donnees/LectureDonnees.cc:12-226: one comment line read with
// getline, then one value read with >>, then ignore() of one character;
// unknown keywords map to 0 (unordered_map::operator[] default insert).
int ref_read_donnees(const char* path, mhy_config* c) {
  ref_config_default(c);
  std::ifstream in(path);
  if (!in) return MHY_ERR_IO;
  static const std::map<std::string, int> cases = {
      {"UnitTestCase", 0},     {"SedovTestCase", 1},   {"TriplePoint", 2},     {"SodCaseX", 4},
      {"SodCaseY", 5},         {"NohTestCase", 6},     {"AdvectionX", 7},      {"AdvectionY", 8},
      {"AdvectionVitX", 9},    {"AdvectionVitY", 10},  {"BiUnitTestCase", 11}, {"BiSedovTestCase", 12},
      {"BiTriplePoint", 13},   {"BiShockBubble", 14},  {"BiSodCaseX", 15},     {"BiSodCaseY", 16},
      {"BiNohTestCase", 17},   {"BiAdvectionX", 18},   {"BiAdvectionY", 19},   {"BiAdvectionVitX", 20},
      {"BiAdvectionVitY", 21}};
  static const std::map<std::string, int> schemas = {{"Eucclhyd", 2000}, {"VNR", 2001}, {"CSTS", 2002}, {"MYR", 2003}};
  static const std::map<std::string, int> lims = {{"minmod", 300},  {"superBee", 301},    {"vanLeer", 302},
                                                  {"minmodG", 1300}, {"superBeeG", 1301}, {"vanLeerG", 1302},
                                                  {"arithmeticG", 1303}, {"ultrabeeG", 1304}};
  static const std::map<std::string, int> yesno = {{"non", 0}, {"oui", 1}};
  static const std::map<std::string, int> dual = {{"A1", 0}, {"A2", 1}, {"PB", 2}};
  static const std::map<std::string, int> eoss = {
      {"Void", 100}, {"PerfectGas", 101}, {"StiffenedGas", 102}, {"Murnhagan", 103}, {"SolidLinear", 104}};
  auto look = [](const std::map<std::string, int>& mp, const std::string& k) {
    auto it = mp.find(k);
    return it == mp.end() ? 0 : it->second;
  };
  std::string line, word;
  std::getline(in, line);
  in >> word;
  c->test_case = look(cases, word);
  in.ignore();
  c->nbmat = (c->test_case < MHY_CASE_BI_UNIT_TEST) ? 1 : 2;
  if (c->test_case == MHY_CASE_BI_TRIPLE_POINT) c->nbmat = 3;
  switch (c->test_case) {
    case 7: case 8: case 18: case 19: case 9: case 10: case 20: case 21: c->sans_lagrange = 1; break;
    default: break;
  }
  std::getline(in, line); in >> c->nx; in.ignore();
  std::getline(in, line); in >> c->ny; in.ignore();
  std::getline(in, line); in >> c->dx; in.ignore();
  std::getline(in, line); in >> c->dy; in.ignore();
  std::getline(in, line); in >> c->final_time; in.ignore();
  std::getline(in, line); in >> c->output_time; in.ignore();
  std::getline(in, line); in >> c->deltat_init; in.ignore();
  std::getline(in, line); in >> c->deltat_min; in.ignore();
  for (int k = 0; k < c->nbmat; ++k) {
    std::getline(in, line); in >> word; c->eos[k] = look(eoss, word); in.ignore();
  }
  std::getline(in, line); in >> word; c->schema = look(schemas, word); in.ignore();
  std::getline(in, line); in >> c->space_order; in.ignore();
  std::getline(in, line); in >> word; c->avec_projection = look(yesno, word); in.ignore();
  std::getline(in, line); in >> word; c->projection_conservative = look(yesno, word); in.ignore();
  std::getline(in, line); in >> c->threshold; in.ignore();
  std::getline(in, line); in >> c->projection_order; in.ignore();
  std::getline(in, line); in >> word; c->limiter = look(lims, word);
  if (c->limiter == -1 || c->limiter == 0) return MHY_ERR_IO;
  in.ignore();
  std::getline(in, line); in >> word; c->pp = look(yesno, word); in.ignore();
  if (c->pp == 1) {
    std::getline(in, line); in >> word; c->pp_mixte = look(yesno, word); in.ignore();
    std::getline(in, line); in >> word; c->limiter_pure = look(lims, word);
    if (c->limiter == -1 && c->pp_mixte == 1) return MHY_ERR_IO;
    in.ignore();
    std::getline(in, line); in >> word; c->pp_complet = look(yesno, word); in.ignore();
  }
  if (c->schema == MHY_SCHEMA_VNR) {
    std::getline(in, line); in >> word; c->pseudo_centree = look(yesno, word); in.ignore();
    if (c->avec_projection == 1) {
      std::getline(in, line); in >> word; c->methode_flux_masse = look(dual, word); in.ignore();
    }
  }
  return MHY_OK;
}

ref_ctx* ref_create(const mhy_config* cfg, int* status) {
  ref_ctx* r = new ref_ctx();
  r->o = *cfg;
  if (cfg->nx < 2 || cfg->ny < 2 || cfg->nbmat < 1 || cfg->nbmat > NMAT) {
    if (status) *status = MHY_ERR_ARG;
    delete r;
    return nullptr;
  }
  r->mesh.nx = cfg->nx, r->mesh.ny = cfg->ny;
  r->mesh.nbCells = (long)cfg->nx * cfg->ny;
  r->mesh.nbNodes = (long)(cfg->nx + 1) * (cfg->ny + 1);
  r->mesh.nbFaces = 2 * r->mesh.nbCells + cfg->nx + cfg->ny;
  r->allocate();
  int st = r->prologue();
  if (status) *status = st;
  if (st != MHY_OK) {
    delete r;
    return nullptr;
  }
  return r;
}

void ref_destroy(ref_ctx* r) { delete r; }

// Single-formula probes for tests/test_oracle_math.py (the oracle's own
// functions, evaluated on given inputs).
double ref_probe_gradient(ref_ctx* r, int lim, double gp, double gm, double phi, double phip, double phim, double h0,
                          double hp, double hm) {
  VE vgp = ezero(), vgm = ezero(), vphi = ezero(), vphip = ezero(), vphim = ezero();
  vgp[0] = gp, vgm[0] = gm, vphi[0] = phi, vphip[0] = phip, vphim[0] = phim;
  return r->computeAndLimitGradPhi(lim, vgp, vgm, vphi, vphip, vphim, h0, hp, hm)[0];
}

double ref_probe_flux_o3(ref_ctx* r, const double* y, const double* h, double vdt) {
  return r->ComputeFluxOrdre3(y[0], y[1], y[2], y[3], y[4], y[5], h[0], h[1], h[2], h[3], h[4], h[5], vdt);
}

double ref_probe_int2y(double X, double x0, double y0, double x1, double y1) {
  return ref_ctx::INT2Y(X, x0, y0, x1, y1);
}

double ref_probe_flux_pp(ref_ctx* r, int lim, double grad, double phi, double phip, double phim, double h0, double hp,
                         double hm, double fnv, double dt, int type) {
  VE g = ezero(), v = ezero(), vp = ezero(), vm = ezero(), flux = ezero(), dual = ezero();
  g[0] = grad, v[0] = phi, vp[0] = phip, vm[0] = phim;
  const int keep = r->o.limiter;
  r->o.limiter = lim;
  r->computeFluxPP(false, g, v, vp, vm, h0, hp, hm, fnv, dt, type, r->o.threshold, 0, 0.0, 0, &flux, &dual);
  r->o.limiter = keep;
  return flux[0];
}

int ref_probe_cell_pp(ref_ctx* r, const double* gp, const double* gm, const double* phi, const double* phip,
                      const double* phim, const int* mix3, const int* pure3, double h0, double hp, double hm,
                      double Fm, double Fp, double fdual, double dt, double* grad, double* ar, double* av,
                      double* dual) {
  VE vgp, vgm, v, vp, vm, g = ezero(), a = ezero(), b = ezero(), d = ezero();
  for (int i = 0; i < NEQ; ++i) vgp[i] = gp[i], vgm[i] = gm[i], v[i] = phi[i], vp[i] = phip[i], vm[i] = phim[i];
  r->gradPhiCell(vgp, vgm, v, vp, vm, mix3, pure3, h0, hp, hm, Fm, Fp, fdual, dt, &g, &a, &b, &d);
  for (int i = 0; i < NEQ; ++i) grad[i] = g[i], ar[i] = a[i], av[i] = b[i], dual[i] = d[i];
  return NEQ;
}

int ref_step(ref_ctx* r) {
  if (r->status != MHY_OK) return r->status;
  bool cont = true;
  int st = r->euc() ? r->eucIterate(false, &cont) : r->iterate(false, &cont);
  if (st != MHY_OK) return r->status = st;
  r->finished = !cont;
  if (r->euc()) r->eucSwap();
  else r->swapAll();
  return MHY_OK;
}

int ref_simulate(ref_ctx* r, const char* output_dir, int max_steps) {
  std::string od = output_dir ? output_dir : "none";
  r->writerDisabled = (od == "none" || od == "NONE");
  if (!r->writerDisabled) {
    r->outdir = od;
    mkdir(od.c_str(), 0755);
  }
  r->n = 0;
  bool cont = true;
  do {
    int st = r->euc() ? r->eucIterate(true, &cont) : r->iterate(true, &cont);
    if (st != MHY_OK) return r->status = st;
    if (max_steps > 0 && r->n >= max_steps) cont = false;
    if (cont) {
      if (r->euc()) r->eucSwap();
      else r->swapAll();
    }
  } while (cont);
  if (!r->euc()) r->dumpVariables();  // Vnr::simulate dumps once more; Eucclhyd::simulate does not
  r->finished = true;
  return MHY_OK;
}

int ref_get_time_state(ref_ctx* r, mhy_time_state* ts) {
  ts->t_n = r->t_n, ts->t_nplus1 = r->t_nplus1;
  ts->deltat_n = r->deltat_n, ts->deltat_nplus1 = r->deltat_nplus1;
  ts->iteration = r->n;
  ts->x_then_y_n = r->x_then_y_n ? 1 : 0;
  ts->finished = r->finished ? 1 : 0;
  ts->status = r->status;
  return MHY_OK;
}

int ref_get_field(ref_ctx* r, int field, double* out, long long n) {
  const long C = r->mesh.nbCells, N = r->mesh.nbNodes;
  auto cellv = [&](const std::vector<double>& v) {
    if (n != C) return MHY_ERR_ARG;
    std::memcpy(out, v.data(), sizeof(double) * C);
    return MHY_OK;
  };
  auto matv = [&](const std::vector<VM>& v, int k) {
    if (n != C) return MHY_ERR_ARG;
    for (long c = 0; c < C; ++c) out[c] = v[c][k];
    return MHY_OK;
  };
  switch (field) {
    case MHY_CELL_DENSITY: return cellv(r->density_n);
    case MHY_CELL_PRESSURE: return cellv(r->pressure_n);
    case MHY_CELL_ENERGY: return cellv(r->energy_n);
    case MHY_CELL_PSEUDO: return cellv(r->pseudo_n);
    case MHY_CELL_SOUND_SPEED: return cellv(r->speed_n);
    case MHY_CELL_DIVU: return cellv(r->divu_n);
    case MHY_CELL_MASS: return cellv(r->cell_mass);
    case MHY_CELL_VELOCITY_X:
    case MHY_CELL_VELOCITY_Y:
      if (!r->euc() || n != C) return MHY_ERR_ARG;
      for (long c = 0; c < C; ++c) out[c] = r->cell_velocity_n[c][field == MHY_CELL_VELOCITY_X ? 0 : 1];
      return MHY_OK;
    case MHY_CELL_OUT_VX:
      if (!r->euc()) return MHY_ERR_ARG;
      return cellv(r->x_cell_velocity);
    case MHY_CELL_OUT_VY:
      if (!r->euc()) return MHY_ERR_ARG;
      return cellv(r->y_cell_velocity);
    case MHY_CELL_MIXTE:
    case MHY_CELL_PURE:
      if (n != C) return MHY_ERR_ARG;
      for (long c = 0; c < C; ++c) out[c] = (field == MHY_CELL_MIXTE) ? r->mixte[c] : r->pure[c];
      return MHY_OK;
    default: break;
  }
  if (field >= MHY_MAT_FRACVOL && field < MHY_MAT_FRACVOL + NMAT * MHY_MAT_STRIDE) {
    int k = (field - MHY_MAT_FRACVOL) / MHY_MAT_STRIDE, f = field - k * MHY_MAT_STRIDE;
    switch (f) {
      case MHY_MAT_FRACVOL: return matv(r->fracvol_env, k);
      case MHY_MAT_MASS_FRACTION: return matv(r->mass_fraction_env, k);
      case MHY_MAT_DENSITY: return matv(r->density_env_n, k);
      case MHY_MAT_ENERGY: return matv(r->energy_env_n, k);
      case MHY_MAT_PRESSURE: return matv(r->pressure_env_n, k);
      case MHY_MAT_PSEUDO: return matv(r->pseudo_env_n, k);
      case MHY_MAT_MASS: return matv(r->cell_mass_env, k);
      case MHY_MAT_SOUND_SPEED: return matv(r->speed_env_n, k);
      default: return MHY_ERR_ARG;
    }
  }
  if (n != N) return MHY_ERR_ARG;
  for (long p = 0; p < N; ++p) {
    switch (field) {
      case MHY_NODE_VELOCITY_X: out[p] = r->node_velocity_n[p][0]; break;
      case MHY_NODE_VELOCITY_Y: out[p] = r->node_velocity_n[p][1]; break;
      case MHY_NODE_MASS: out[p] = r->node_mass[p]; break;
      case MHY_NODE_COORD_X: out[p] = r->node_coord_n[p][0]; break;
      case MHY_NODE_COORD_Y: out[p] = r->node_coord_n[p][1]; break;
      case MHY_NODE_OUT_VX: out[p] = r->x_velocity[p]; break;
      case MHY_NODE_OUT_VY: out[p] = r->y_velocity[p]; break;
      default: return MHY_ERR_ARG;
    }
  }
  return MHY_OK;
}

int ref_set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
  return omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

}  // extern "C"
