// Context of one problem on one GPU: device allocation and upload of the
// initial state, the per-step kernel sequence on the context stream, field
// access, conservation sums, kernel timing, and the extern "C" ABI declared
// in include/mahyco.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "host/init.hpp"
#include "host/output.hpp"
#include "kernels.hpp"

using namespace mhy;

namespace {

enum KernelId {
  K_DT = 0, K_NODE, K_CELL, K_SWEEP_X, K_SWEEP_Y, K_DUAL_X, K_DUAL_Y, K_FINAL, K_LAGSWEEP_X, K_LAGSWEEP_Y, K_COUNT
};
const char* kKernelNames[K_COUNT] = {"dt_reduce", "lagrange_node", "lagrange_cell",    "sweep_x",
                                     "sweep_y",   "dual_x",        "dual_y",           "remap_final",
                                     "lagrange_sweep_x", "lagrange_sweep_y"};

constexpr int kBlock = 256;
constexpr int kReduceBlocks = 1024;

// Block partial sums of the conservation reduction: at most
// kConsBlocks blocks x (2 + MHY_NBMATMAX) quantities x (sum, compensation)
// x 2 passes.
constexpr int kConsBlocks = 256;
constexpr int kConsScratch = kConsBlocks * (2 + MHY_NBMATMAX) * 4;  // doubles (40 KiB)

}  // namespace

// Scratch of mhy_conservation_noalloc: one copy per GPU, held until the
// process ends.  Host code never passes this name to a kernel: the address
// of each device's copy comes from hipGetSymbolAddress at context set-up
// (scratchAddress) and is what the kernels receive.
static __device__ __align__(256) double scratch[kConsScratch];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratchLock;
double* g_scratchAddr[kMaxDevices] = {nullptr};

// Looks up this device's copy of the scratch array once per device (the
// device must be current); returns nullptr on failure.
double* scratchAddress(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratchLock);
  if (!g_scratchAddr[device]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    g_scratchAddr[device] = static_cast<double*>(ptr);
  }
  return g_scratchAddr[device];
}

}  // namespace

struct mhy_ctx {
  mhy_config cfg;
  Opts o;
  // device pointers; in the mixed form the bulk fields (FieldsT<float>
  // members of type R) point to FP32 arrays and are retyped at launch
  Fields f;
  bool mixed = false;  // precision 1: FP32 bulk storage and sweep arithmetic
  // the Eucclhyd scheme (eucclhyd.hip): its own Lagrange phase and final
  // remap around the cell sweeps, no nodal remap
  bool euc = false;
  EucFields ef;
  HostInit init;
  hipStream_t stream = nullptr;
  int device = 0;
  int cur = 0;
  bool x_then_y_n = true;
  int iteration = 0;
  bool finished = false;
  // sweep and dual kernels: wave-shuffle (default) or per-cell/per-node
  // stencil (environment MAHYCO_SWEEP_KERNEL=cell selects the latter)
  bool sweepWave = true;
  // sweep 2 runs the cell part of the final remap (environment
  // MAHYCO_FUSE_FINAL=0 keeps it a separate kernel); finalDone records that
  // the current iteration's final remap is already done
  bool fuseFinal = true, finalDone = false;
  // sweep 1 runs the cell part of the Lagrange phase (wave sweeps only;
  // environment MAHYCO_FUSE_LAGRANGE=0 keeps it a separate kernel)
  bool fuseLag = true;
  // ... also on steps whose first sweep runs along y (the tuner may turn it
  // off there: the y strips recompute a third of their cells as halo)
  // (MAHYCO_LAG_Y=fused / separate forces either)
  bool fuseLagY = true;
  // with the fused first sweep: that sweep stores the Lagrange volume and
  // centroid for the second (environment MAHYCO_SWEEP_GEOM=load), or the
  // second sweep forms them again from the node velocities (=recompute)
  bool geomStored = true;
  // node update on the Euler grid as an interior kernel and a boundary
  // kernel (MAHYCO_SPLIT_NODE=0: one kernel for all nodes)
  bool splitNode = true;
  // nodal remap along y: the wave kernel (default) or the per-node stencil
  // kernel (environment MAHYCO_DUAL_Y=node)
  bool dualWaveY = true;
  // nodal remap along y with the column-walking kernel: chosen for the
  // mixed form on meshes of at least 8192 walking waves, where it measured
  // faster (fewer waves leave the chip underfilled); MAHYCO_DUAL_Y=walk /
  // wave / node forces a kernel
  bool dualWalkY = false;
  // nodal remap along y on LDS tiles of 16 node columns x dualTileRows
  // rows: 48 (MAHYCO_DUAL_Y=tile), 32 or 64 (tile32, tile64)
  bool dualTileY = false;
  int dualTileRows = 48;
  // cell sweeps along y: the column-walking kernel or the wave-strip
  // kernel.  The walker keeps more state in registers (fewer waves per SIMD)
  // but has less halo work; it is chosen for the mixed form with one
  // material and limited slopes on meshes of at least 4096 walking waves,
  // where it measured faster.  MAHYCO_SWEEP_Y=walk / wave forces either.
  bool walkY = false;
  // mhy_run replays each step as a captured HIP graph, one per step parity
  // and sweep order (index 2 * cur + x_then_y_n); MAHYCO_GRAPH=0 launches
  // eagerly
  bool useGraph = true;
  hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
  int status = MHY_OK;
  // this device's copy of the scratch array (mhy_conservation_noalloc)
  double* scratch = nullptr;
  long ncell = 0, nnode = 0, nface = 0;
  std::string err;
  std::vector<void*> allocs;
  // kernel timing
  bool timing = false;
  unsigned timingMask = ~0u;  // kernels (bit 1 << KernelId) timed when timing is on
  struct Rec {
    int k;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> eventPool;
  double ms[K_COUNT] = {0};
  int launches[K_COUNT] = {0};
  // Launch-shape tuner.  The y-direction kernels come in interchangeable
  // variants with different tilings (results bitwise the same): the nodal
  // remap as 16x64-node LDS tiles, 4x16-node wave strips or 64-column
  // walking bands; the cell sweep as 4x16-cell wave strips or 64-column
  // walking bands; and, on steps whose first sweep runs along y, the
  // Lagrange cell update fused into that sweep or run as its own kernel.  On
  // a large problem the first steps time whole steps, one site after the
  // other (the other sites keep their current choice): kTuneRounds rounds
  // over the site's candidates, two steps per candidate and round (one of
  // each sweep order), and keep the candidate with the fastest round, so
  // that clock ramp-up in the first steps and single slow steps do not
  // decide.  The choice is fixed before any graph is captured.
  static constexpr int kTuneRounds = 2;
  static constexpr double kTuneMargin = 0.01;  // relative gain a non-default candidate must show
  struct Tuner {
    bool on = false;
    int step = 0, steps = 0;
    static constexpr int kMaxCand = 5;
    int ncand[3] = {0, 0, 0};  // site 0: nodal remap along y, 1: cell sweep along y, 2: y-first Lagrange
    int cand[3][kMaxCand] = {};
    int best[3] = {0, 0, 0};  // index into cand of the current choice
    double ms[3][kTuneRounds][kMaxCand] = {};  // two-step time per site, round and candidate
    hipEvent_t ev[2] = {nullptr, nullptr};  // around the current step
    std::string choice;  // the kept variants, for reports
    std::string forcedLag;  // the y-first Lagrange kernel forced through MAHYCO_LAG_Y
  } tune;
  // output
  OutputWriter* writer = nullptr;
};

namespace {

int fail(mhy_ctx* x, int code, const std::string& msg) {
  if (x) {
    x->err = msg;
    x->status = code;
  }
  return code;
}

#define HIPCHK(x, call)                                                                             \
  do {                                                                                              \
    hipError_t e_ = (call);                                                                         \
    if (e_ != hipSuccess) return fail((x), MHY_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Allocation helpers keep the first error in e and skip later requests.
double* allocD(mhy_ctx* x, long n, hipError_t& e) {
  if (e != hipSuccess) return nullptr;
  void* p = nullptr;
  e = hipMalloc(&p, sizeof(double) * (size_t)std::max(n, 1L));
  if (e == hipSuccess) {
    x->allocs.push_back(p);
    e = hipMemset(p, 0, sizeof(double) * (size_t)std::max(n, 1L));
  }
  return static_cast<double*>(p);
}

// bulk-field array: FP64, or FP32 in the mixed form (kept as an untyped
// pointer in Fields)
double* allocR(mhy_ctx* x, long n, hipError_t& e) {
  if (e != hipSuccess) return nullptr;
  void* p = nullptr;
  const size_t bytes = (x->mixed ? sizeof(float) : sizeof(double)) * (size_t)std::max(n, 1L);
  e = hipMalloc(&p, bytes);
  if (e == hipSuccess) {
    x->allocs.push_back(p);
    e = hipMemset(p, 0, bytes);
  }
  return static_cast<double*>(p);
}

int* allocI(mhy_ctx* x, long n, hipError_t& e) {
  if (e != hipSuccess) return nullptr;
  void* p = nullptr;
  e = hipMalloc(&p, sizeof(int) * (size_t)std::max(n, 1L));
  if (e == hipSuccess) {
    x->allocs.push_back(p);
    e = hipMemset(p, 0, sizeof(int) * (size_t)std::max(n, 1L));
  }
  return static_cast<int*>(p);
}

KP makeKP(const mhy_ctx* x) {
  KP kp;
  kp.f = x->f;
  kp.o = x->o;
  kp.cur = x->cur;
  kp.nxt = 1 - x->cur;
  kp.sweep = 1;
  kp.xdir = 1;
  kp.fuseFinal = 0;
  // the stored Lagrange geometry exists unless the fused first sweep keeps
  // it in registers only
  kp.geomLoad = (!x->fuseLag || x->geomStored) ? 1 : 0;
  kp.ncell = x->ncell;
  kp.nnode = x->nnode;
  return kp;
}

// candidate ids of the tuner's sites
enum { TD_TILE = 0, TD_WAVE = 1, TD_WALK = 2, TD_TILE32 = 3, TD_TILE64 = 4 };  // nodal remap along y
enum { TS_WAVE = 0, TS_WALK = 1 };               // cell sweep along y
enum { TL_FUSED = 0, TL_SEPARATE = 1 };          // Lagrange cell update of y-first steps
const char* kDualNames[5] = {"tile", "wave", "walk", "tile32", "tile64"};
const char* kSweepNames[2] = {"wave", "walk"};
const char* kLagNames[2] = {"fused", "separate"};

void setDualY(mhy_ctx* x, int c) {
  x->dualTileY = (c == TD_TILE || c == TD_TILE32 || c == TD_TILE64);
  x->dualTileRows = (c == TD_TILE32) ? 32 : (c == TD_TILE64) ? 64 : 48;
  x->dualWalkY = (c == TD_WALK);
  x->dualWaveY = true;
}

// Applies candidate index ci of a site.
void tuneApply(mhy_ctx* x, int site, int ci) {
  const int id = x->tune.cand[site][ci];
  if (site == 0) setDualY(x, id);
  else if (site == 1) x->walkY = (id == TS_WALK);
  else x->fuseLagY = (id == TL_FUSED);
}

// The site, round and candidate index of tuning step k (sites one after the
// other; within a site, rounds over the candidates, two steps each); false
// past the schedule.
bool tuneSlot(const mhy_ctx* x, int k, int& site, int& ci, int& round) {
  const auto& t = x->tune;
  for (site = 0; site < 3; ++site) {
    const int per = 2 * t.ncand[site];
    if (k < mhy_ctx::kTuneRounds * per) {
      round = k / per;
      ci = (k % per) / 2;
      return true;
    }
    k -= mhy_ctx::kTuneRounds * per;
  }
  return false;
}

// Selects the candidate of the current tuning step and starts its timing
// (called at the start of every step while tuning).
void tuneSelect(mhy_ctx* x) {
  auto& t = x->tune;
  if (!t.on) return;
  int site, ci, round;
  if (!tuneSlot(x, t.step, site, ci, round)) return;
  tuneApply(x, site, ci);
  (void)hipEventRecord(t.ev[0], x->stream);
}

// Ends a tuning step: adds its time to its candidate's round; after the
// last step of a site, keeps the candidate with the fastest round.
void tuneStepDone(mhy_ctx* x) {
  auto& t = x->tune;
  if (!t.on) return;
  int site, ci, round;
  if (!tuneSlot(x, t.step, site, ci, round)) return;
  (void)hipEventRecord(t.ev[1], x->stream);
  float ms = 0.f;
  if (hipEventSynchronize(t.ev[1]) == hipSuccess && hipEventElapsedTime(&ms, t.ev[0], t.ev[1]) == hipSuccess)
    t.ms[site][round][ci] += ms;
  ++t.step;
  int nsite, nci, nround;
  const bool more = tuneSlot(x, t.step, nsite, nci, nround);
  if (!more || nsite != site) {  // the site is done
    auto fastest = [&](int c) {
      double m = t.ms[site][0][c];
      for (int r = 1; r < mhy_ctx::kTuneRounds; ++r) m = std::min(m, t.ms[site][r][c]);
      return m;
    };
    // another candidate replaces the default (index 0) only for a clear
    // gain: candidates within the step-to-step noise keep the default, so
    // that repeated runs keep the same kernels
    int b = 0;
    for (int c = 1; c < t.ncand[site]; ++c)
      if (fastest(c) < (1.0 - mhy_ctx::kTuneMargin) * fastest(b)) b = c;
    t.best[site] = b;
    tuneApply(x, site, b);
  }
  if (!more) {
    t.on = false;
    t.choice.clear();
    auto add = [&](const char* k, const char* v) {
      t.choice += std::string(t.choice.empty() ? "" : " ") + k + "=" + v;
    };
    if (t.ncand[0]) add("dual_y", kDualNames[t.cand[0][t.best[0]]]);
    if (t.ncand[1]) add("sweep_y", kSweepNames[t.cand[1][t.best[1]]]);
    if (t.ncand[2]) add("lagrange_y", kLagNames[t.cand[2][t.best[2]]]);
    else if (!t.forcedLag.empty()) add("lagrange_y", t.forcedLag.c_str());
  }
}

hipEvent_t takeEvent(mhy_ctx* x) {
  if (!x->eventPool.empty()) {
    hipEvent_t e = x->eventPool.back();
    x->eventPool.pop_back();
    return e;
  }
  hipEvent_t e;
  (void)hipEventCreate(&e);
  return e;
}

// Launch helper: optional HIP events around the kernel on the context stream.
template <class F>
void timed(mhy_ctx* x, int k, F&& launch) {
  if (!x->timing || !((x->timingMask >> k) & 1u)) {
    launch();
    return;
  }
  mhy_ctx::Rec r;
  r.k = k;
  r.a = takeEvent(x);
  r.b = takeEvent(x);
  (void)hipEventRecord(r.a, x->stream);
  launch();
  (void)hipEventRecord(r.b, x->stream);
  x->pending.push_back(r);
}

void collectTimes(mhy_ctx* x) {
  for (auto& r : x->pending) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
      x->ms[r.k] += t;
      x->launches[r.k] += 1;
    }
    x->eventPool.push_back(r.a);
    x->eventPool.push_back(r.b);
  }
  x->pending.clear();
}

dim3 gridFor(long n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

// launches kd (FP64 form) or kf (mixed form) on the context stream with the
// parameter block seen with the matching storage type
template <class KD, class KF, class... A>
void launchR(const mhy_ctx* x, KD kd, KF kf, dim3 grid, dim3 block, const KP& kp, A... args) {
  if (x->mixed) hipLaunchKernelGGL(kf, grid, block, 0, x->stream, retype<float>(kp), args...);
  else hipLaunchKernelGGL(kd, grid, block, 0, x->stream, kp, args...);
}

int checkLaunch(mhy_ctx* x, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(x, MHY_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  return MHY_OK;
}

// ------------------------------------------------ Eucclhyd phases
using EucKernelFn = void (*)(KP, EucFields);
using EucSweepFn = void (*)(KP);

// State-of-step kernel and the time step (computeEOS .. computedeltat).
int phaseEucStart(mhy_ctx* x) {
  KP kp = makeKP(x);
  static const EucKernelFn ks[3] = {k_euc_start<1>, k_euc_start<2>, k_euc_start<3>};
  const unsigned nb = (unsigned)std::min<long>(kReduceBlocks, (x->ncell + kBlock - 1) / kBlock);
  timed(x, K_DT, [&] {
    hipLaunchKernelGGL(ks[x->o.nbmat - 1], dim3(nb), dim3(kBlock), 0, x->stream, kp, x->ef);
    hipLaunchKernelGGL(k_euc_dt_finalize, dim3(1), dim3(64), 0, x->stream, kp);
  });
  return checkLaunch(x, "eucclhyd start");
}

// Node solver and cell update (extrapolateValue ..
// updateCellCenteredLagrangeVariables).
int phaseEucLagrange(mhy_ctx* x) {
  KP kp = makeKP(x);
  static const EucKernelFn kn[3] = {k_euc_node<1>, k_euc_node<2>, k_euc_node<3>};
  static const EucKernelFn kc[3] = {k_euc_cell<1>, k_euc_cell<2>, k_euc_cell<3>};
  const int mi = x->o.nbmat - 1;
  timed(x, K_NODE, [&] { hipLaunchKernelGGL(kn[mi], gridFor(x->nnode), dim3(kBlock), 0, x->stream, kp, x->ef); });
  timed(x, K_CELL, [&] { hipLaunchKernelGGL(kc[mi], gridFor(x->ncell), dim3(kBlock), 0, x->stream, kp, x->ef); });
  return checkLaunch(x, "eucclhyd lagrange");
}

// remapCellcenteredVariable
int phaseEucFinal(mhy_ctx* x) {
  KP kp = makeKP(x);
  static const EucKernelFn kf[3] = {k_euc_final<1>, k_euc_final<2>, k_euc_final<3>};
  timed(x, K_FINAL, [&] {
    hipLaunchKernelGGL(kf[x->o.nbmat - 1], gridFor(x->ncell), dim3(kBlock), 0, x->stream, kp, x->ef);
  });
  return checkLaunch(x, "eucclhyd final");
}

// The cell sweeps on the Eucclhyd vector layout (4 + m, or 8 + m with the
// kinetic energy of a conservative projection): wave strips, or the LDS
// tile kernel for the third order.
template <int L>
EucSweepFn eucSweepFor(bool xd, bool pp, bool o3) {
  if (o3) return xd ? k_sweep_o3<double, L, true> : k_sweep_o3<double, L, false>;
  if (pp) return xd ? k_sweep_wave<double, L, true, true> : k_sweep_wave<double, L, true, false>;
  return xd ? k_sweep_wave<double, L, false, true> : k_sweep_wave<double, L, false, false>;
}
EucSweepFn eucSweepKernel(bool xd, int m, bool pp, bool conservative, bool o3) {
  switch (m + (conservative ? 8 : 4)) {
    case 5: return eucSweepFor<5>(xd, pp, o3);
    case 6: return eucSweepFor<6>(xd, pp, o3);
    case 7: return eucSweepFor<7>(xd, pp, o3);
    case 9: return eucSweepFor<9>(xd, pp, o3);
    case 10: return eucSweepFor<10>(xd, pp, o3);
    default: return eucSweepFor<11>(xd, pp, o3);
  }
}

int phaseEucSweep(mhy_ctx* x, int sweep, int stage) {
  if (stage != MHY_STAGE_CELLS) return MHY_OK;  // no nodal remap
  KP kp = makeKP(x);
  kp.sweep = sweep;
  const bool xd = (sweep == 1) ? x->x_then_y_n : !x->x_then_y_n;
  kp.xdir = xd ? 1 : 0;
  const bool o3 = (x->o.order == 3), pp = (x->o.pp == 1);
  const dim3 grid = o3 ? sweepO3Grid(xd, x->o.nx, x->o.ny) : sweepWaveGrid(xd, x->o.nx, x->o.ny, false, pp);
  const dim3 block(o3 ? kBlock : sweepWaveBlock(xd, false, pp));
  const EucSweepFn k = eucSweepKernel(xd, x->o.nbmat, pp, x->o.conservative == 1, o3);
  timed(x, xd ? K_SWEEP_X : K_SWEEP_Y, [&] { hipLaunchKernelGGL(k, grid, block, 0, x->stream, kp); });
  return checkLaunch(x, "eucclhyd sweep");
}

int phaseDt(mhy_ctx* x) {
  if (x->euc) return phaseEucStart(x);
  tuneSelect(x);  // a step starts here
  KP kp = makeKP(x);
  if (x->o.mode == MODE_ADVECTION) {
    timed(x, K_DT, [&] { launchR(x, k_dt_hold<double>, k_dt_hold<float>, dim3(1), dim3(64), kp); });
  } else {
    const unsigned nb = (unsigned)std::min<long>(kReduceBlocks, (x->ncell + kBlock - 1) / kBlock);
    timed(x, K_DT, [&] {
      launchR(x, k_dt_reduce<double>, k_dt_reduce<float>, dim3(nb), dim3(kBlock), kp);
      launchR(x, k_dt_finalize<double>, k_dt_finalize<float>, dim3(1), dim3(64), kp);
    });
  }
  return checkLaunch(x, "dt");
}

using KernelFn = void (*)(KP);
using KernelFnF = void (*)(KPT<float>);

// the Lagrange cell update of this step runs inside the first sweep
bool lagFused(const mhy_ctx* x) { return x->fuseLag && (x->x_then_y_n || x->fuseLagY); }

int phaseLagrange(mhy_ctx* x) {
  if (x->euc) return phaseEucLagrange(x);
  KP kp = makeKP(x);
  // (a grid's y dimension holds at most 65535 node rows)
  if (x->o.mode == MODE_FULL && x->o.nx >= 2 && x->o.ny >= 2 && x->o.ny - 1 <= 65535 && x->splitNode) {
    const dim3 gi((unsigned)((x->o.nx - 1 + kBlock - 1) / kBlock), (unsigned)(x->o.ny - 1));
    const long nedge = 2L * (x->o.nx + 1) + 2L * (x->o.ny - 1);
    timed(x, K_NODE, [&] {
      launchR(x, k_lagrange_node_inner<double>, k_lagrange_node_inner<float>, gi, dim3(kBlock), kp);
      launchR(x, k_lagrange_node_edge<double>, k_lagrange_node_edge<float>, gridFor(nedge), dim3(kBlock), kp);
    });
  } else {
    timed(x, K_NODE, [&] {
      launchR(x, k_lagrange_node<double>, k_lagrange_node<float>, gridFor(x->nnode), dim3(kBlock), kp);
    });
  }
  static const KernelFn kc[3] = {k_lagrange_cell<double, 1>, k_lagrange_cell<double, 2>, k_lagrange_cell<double, 3>};
  static const KernelFnF kcf[3] = {k_lagrange_cell<float, 1>, k_lagrange_cell<float, 2>, k_lagrange_cell<float, 3>};
  const int mi = x->o.nbmat - 1;
  // with the fused first sweep the cell update runs in that sweep
  if (!(lagFused(x) && x->o.mode != MODE_LAGRANGE))
    timed(x, K_CELL, [&] { launchR(x, kc[mi], kcf[mi], gridFor(x->ncell), dim3(kBlock), kp); });
  return checkLaunch(x, "lagrange");
}

// sweep kernels for the direction, the material count and the PenteBorne
// option: the wave-shuffle kernels (default) or the per-cell stencil kernels
// (FP64 form only)
template <class R>
void (*sweepWaveKernel(bool xd, int m, bool pp, bool lag = false))(KPT<R>) {
  using Fn = void (*)(KPT<R>);
  static const Fn kl[2][3][2] = {
      {{k_sweep_wave_lag<R, 1, false, false>, k_sweep_wave_lag<R, 1, true, false>},
       {k_sweep_wave_lag<R, 2, false, false>, k_sweep_wave_lag<R, 2, true, false>},
       {k_sweep_wave_lag<R, 3, false, false>, k_sweep_wave_lag<R, 3, true, false>}},
      {{k_sweep_wave_lag<R, 1, false, true>, k_sweep_wave_lag<R, 1, true, true>},
       {k_sweep_wave_lag<R, 2, false, true>, k_sweep_wave_lag<R, 2, true, true>},
       {k_sweep_wave_lag<R, 3, false, true>, k_sweep_wave_lag<R, 3, true, true>}}};
  if (lag) return kl[xd ? 1 : 0][m - 1][pp ? 1 : 0];
  static const Fn kw[2][3][2] = {
      {{k_sweep_wave<R, 1, false, false>, k_sweep_wave<R, 1, true, false>},
       {k_sweep_wave<R, 2, false, false>, k_sweep_wave<R, 2, true, false>},
       {k_sweep_wave<R, 3, false, false>, k_sweep_wave<R, 3, true, false>}},
      {{k_sweep_wave<R, 1, false, true>, k_sweep_wave<R, 1, true, true>},
       {k_sweep_wave<R, 2, false, true>, k_sweep_wave<R, 2, true, true>},
       {k_sweep_wave<R, 3, false, true>, k_sweep_wave<R, 3, true, true>}}};
  return kw[xd ? 1 : 0][m - 1][pp ? 1 : 0];
}
KernelFn sweepKernel(bool wave, bool xd, int m, bool pp) {
  if (wave) return sweepWaveKernel<double>(xd, m, pp);
  static const KernelFn kc[2][3][2] = {
      {{k_sweep_cell<1, false, false>, k_sweep_cell<1, true, false>},
       {k_sweep_cell<2, false, false>, k_sweep_cell<2, true, false>},
       {k_sweep_cell<3, false, false>, k_sweep_cell<3, true, false>}},
      {{k_sweep_cell<1, false, true>, k_sweep_cell<1, true, true>},
       {k_sweep_cell<2, false, true>, k_sweep_cell<2, true, true>},
       {k_sweep_cell<3, false, true>, k_sweep_cell<3, true, true>}}};
  return kc[xd ? 1 : 0][m - 1][pp ? 1 : 0];
}

// One remap sweep: stage 1 = cell fluxes and update (computeGradPhiFace ..
// computeUremap), stage 2 = nodal momentum remap (computeDualUremap).
int phaseSweepStage(mhy_ctx* x, int sweep, int stage) {
  if (x->euc) return phaseEucSweep(x, sweep, stage);
  KP kp = makeKP(x);
  kp.sweep = sweep;
  if (stage == MHY_STAGE_CELLS) {
    // the cell stage of sweep 2 also runs the cell part of remapVariables
    kp.fuseFinal = (sweep == 2 && x->fuseFinal) ? 1 : 0;
    x->finalDone = (kp.fuseFinal == 1);
  }
  const bool xd = (sweep == 1) ? x->x_then_y_n : !x->x_then_y_n;
  kp.xdir = xd ? 1 : 0;
  if (stage == MHY_STAGE_CELLS) {
    // third order: LDS-tile kernel (the nodal stage keeps its own choice)
    const bool o3 = (x->o.order == 3), wave = x->sweepWave && !o3;
    const dim3 grid = wave ? sweepWaveGrid(xd, x->o.nx, x->o.ny, x->mixed, x->o.pp == 1)
                      : o3 ? sweepO3Grid(xd, x->o.nx, x->o.ny)
                           : gridFor(x->ncell);
    const dim3 block(wave ? sweepWaveBlock(xd, x->mixed, x->o.pp == 1) : kBlock);
    const bool lagSweep = (sweep == 1 && lagFused(x));
    if (!xd && wave && x->walkY && !lagSweep) {
      static const KernelFn kd[3][2] = {{k_sweep_walk<double, 1, false>, k_sweep_walk<double, 1, true>},
                                        {k_sweep_walk<double, 2, false>, k_sweep_walk<double, 2, true>},
                                        {k_sweep_walk<double, 3, false>, k_sweep_walk<double, 3, true>}};
      static const KernelFnF kf[3][2] = {{k_sweep_walk<float, 1, false>, k_sweep_walk<float, 1, true>},
                                         {k_sweep_walk<float, 2, false>, k_sweep_walk<float, 2, true>},
                                         {k_sweep_walk<float, 3, false>, k_sweep_walk<float, 3, true>}};
      const int mi = x->o.nbmat - 1, pi = x->o.pp == 1 ? 1 : 0;
      timed(x, K_SWEEP_Y, [&] {
        launchR(x, kd[mi][pi], kf[mi][pi], sweepWalkGrid(x->o.nx, x->o.ny), dim3(kBlock), kp);
      });
    } else if (o3) {  // third order: LDS-tile kernel, no PenteBorne (checked at creation)
      static const KernelFn kd[2][3] = {
          {k_sweep_o3<double, 1, false>, k_sweep_o3<double, 2, false>, k_sweep_o3<double, 3, false>},
          {k_sweep_o3<double, 1, true>, k_sweep_o3<double, 2, true>, k_sweep_o3<double, 3, true>}};
      static const KernelFnF kf[2][3] = {
          {k_sweep_o3<float, 1, false>, k_sweep_o3<float, 2, false>, k_sweep_o3<float, 3, false>},
          {k_sweep_o3<float, 1, true>, k_sweep_o3<float, 2, true>, k_sweep_o3<float, 3, true>}};
      const int di = xd ? 1 : 0, mi = x->o.nbmat - 1;
      timed(x, xd ? K_SWEEP_X : K_SWEEP_Y, [&] { launchR(x, kd[di][mi], kf[di][mi], grid, block, kp); });
    } else if (lagSweep) {  // with the cell part of the Lagrange phase
      const KernelId kid = xd ? K_LAGSWEEP_X : K_LAGSWEEP_Y;
      if (x->mixed) {
        const KernelFnF k = sweepWaveKernel<float>(xd, x->o.nbmat, x->o.pp == 1, true);
        timed(x, kid, [&] { hipLaunchKernelGGL(k, grid, block, 0, x->stream, retype<float>(kp)); });
      } else {
        const KernelFn k = sweepWaveKernel<double>(xd, x->o.nbmat, x->o.pp == 1, true);
        timed(x, kid, [&] { hipLaunchKernelGGL(k, grid, block, 0, x->stream, kp); });
      }
    } else if (x->mixed) {
      const KernelFnF k = sweepWaveKernel<float>(xd, x->o.nbmat, x->o.pp == 1);
      timed(x, xd ? K_SWEEP_X : K_SWEEP_Y,
            [&] { hipLaunchKernelGGL(k, grid, block, 0, x->stream, retype<float>(kp)); });
    } else {
      const KernelFn k = sweepKernel(wave, xd, x->o.nbmat, x->o.pp == 1);
      timed(x, xd ? K_SWEEP_X : K_SWEEP_Y, [&] { hipLaunchKernelGGL(k, grid, block, 0, x->stream, kp); });
    }
  } else if (x->sweepWave && xd) {
    const dim3 grid = dualWaveGrid(true, x->o.nx, x->o.ny, x->mixed);
    timed(x, K_DUAL_X, [&] { launchR(x, k_dual_wave<double, true>, k_dual_wave<float, true>, grid, dim3(kBlock), kp); });
  } else if (xd) {
    timed(x, K_DUAL_X,
          [&] { launchR(x, k_dual<double, true>, k_dual<float, true>, gridFor(x->nnode), dim3(kBlock), kp); });
  } else if (x->dualTileY) {
    timed(x, K_DUAL_Y, [&] {
      const int rows = x->dualTileRows;
      const dim3 grid = dualTileGrid(x->o.nx, x->o.ny, rows), block(dualTileBlock());
      if (rows == 32) launchR(x, k_dual_tile<double, 32>, k_dual_tile<float, 32>, grid, block, kp);
      else if (rows == 64) launchR(x, k_dual_tile<double, 64>, k_dual_tile<float, 64>, grid, block, kp);
      else launchR(x, k_dual_tile<double, 48>, k_dual_tile<float, 48>, grid, block, kp);
    });
  } else if (x->dualWalkY) {
    timed(x, K_DUAL_Y, [&] {
      launchR(x, k_dual_walk<double>, k_dual_walk<float>, dualWalkGrid(x->o.nx, x->o.ny), dim3(kBlock), kp);
    });
  } else if (x->sweepWave && x->dualWaveY) {
    const dim3 grid = dualWaveGrid(false, x->o.nx, x->o.ny, x->mixed);
    timed(x, K_DUAL_Y, [&] {
      launchR(x, k_dual_wave<double, false>, k_dual_wave<float, false>, grid, dim3(dualWaveBlock(false, x->mixed)),
              kp);
    });
  } else {
    timed(x, K_DUAL_Y,
          [&] { launchR(x, k_dual<double, false>, k_dual<float, false>, gridFor(x->nnode), dim3(kBlock), kp); });
  }
  return checkLaunch(x, "sweep");
}

int phaseSweep(mhy_ctx* x, int sweep) {
  int st = phaseSweepStage(x, sweep, MHY_STAGE_CELLS);
  return st != MHY_OK ? st : phaseSweepStage(x, sweep, MHY_STAGE_NODES);
}

int phaseFinal(mhy_ctx* x) {
  if (x->euc) return phaseEucFinal(x);
  if (x->finalDone) {  // already done by the cell stage of sweep 2
    x->finalDone = false;
    return MHY_OK;
  }
  KP kp = makeKP(x);
  static const KernelFn kf[3] = {k_remap_final<double, 1>, k_remap_final<double, 2>, k_remap_final<double, 3>};
  static const KernelFnF kff[3] = {k_remap_final<float, 1>, k_remap_final<float, 2>, k_remap_final<float, 3>};
  const int m = x->o.nbmat - 1;
  timed(x, K_FINAL, [&] { launchR(x, kf[m], kff[m], gridFor(x->ncell), dim3(kBlock), kp); });
  return checkLaunch(x, "final");
}

int readScalars(mhy_ctx* x, Scalars& s) {
  HIPCHK(x, hipMemcpyAsync(&s, x->f.sc, sizeof(Scalars), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(x, hipStreamSynchronize(x->stream));
  collectTimes(x);
  return MHY_OK;
}

// Swap of the n / n+1 buffers (VnrRemap.cc:198-222) = parity flip.
void endIteration(mhy_ctx* x) {
  tuneStepDone(x);
  x->cur = 1 - x->cur;
  if (x->o.mode != MODE_LAGRANGE) x->x_then_y_n = !x->x_then_y_n;
}

int runIteration(mhy_ctx* x) {
  int st;
  if ((st = phaseDt(x)) != MHY_OK) return st;
  if ((st = phaseLagrange(x)) != MHY_OK) return st;
  if (x->o.mode != MODE_LAGRANGE) {
    if ((st = phaseSweep(x, 1)) != MHY_OK) return st;
    if ((st = phaseSweep(x, 2)) != MHY_OK) return st;
    if ((st = phaseFinal(x)) != MHY_OK) return st;
  }
  x->iteration++;
  return MHY_OK;
}

int upload(mhy_ctx* x, double* dst, const std::vector<double>& src) {
  HIPCHK(x, hipMemcpy(dst, src.data(), sizeof(double) * src.size(), hipMemcpyHostToDevice));
  return MHY_OK;
}

// upload of a bulk field (rounded to FP32 in the mixed form)
int uploadR(mhy_ctx* x, double* dst, const std::vector<double>& src) {
  if (!x->mixed) return upload(x, dst, src);
  std::vector<float> h(src.begin(), src.end());
  HIPCHK(x, hipMemcpy(dst, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
  return MHY_OK;
}

// true for the fields stored with the bulk type R (FP32 in the mixed form)
bool isBulkField(int field) {
  switch (field) {
    case MHY_CELL_MASS:
    case MHY_NODE_COORD_X:
    case MHY_NODE_COORD_Y:
      return false;
    default: break;
  }
  if (field >= MHY_MAT_FRACVOL && field < MHY_MAT_FRACVOL + MHY_NBMATMAX * MHY_MAT_STRIDE) {
    const int id = field - ((field - MHY_MAT_FRACVOL) / MHY_MAT_STRIDE) * MHY_MAT_STRIDE;
    return id != MHY_MAT_FRACVOL && id != MHY_MAT_MASS;
  }
  return true;
}

bool isMaterialField(int field) {
  return field >= MHY_MAT_FRACVOL && field < MHY_MAT_FRACVOL + MHY_NBMATMAX * MHY_MAT_STRIDE;
}

// compact1: writes the derived material fields of the cells that do not
// store them into the material arrays (the flags stay), so that the arrays
// hold every cell's values
int expandMaterialFields(mhy_ctx* x) {
  if (!x->o.compact1) return MHY_OK;
  KP kp = makeKP(x);
  launchR(x, k_env_expand<double>, k_env_expand<float>, gridFor(x->ncell), dim3(kBlock), kp);
  HIPCHK(x, hipGetLastError());
  return MHY_OK;
}

// compact1: expands the material fields and marks every cell as storing them
int storeMaterialFields(mhy_ctx* x) {
  if (!x->o.compact1) return MHY_OK;
  const int st = expandMaterialFields(x);
  if (st != MHY_OK) return st;
  HIPCHK(x, hipMemsetAsync(x->f.envx, 1, x->ncell, x->stream));
  return MHY_OK;
}

// device pointer and length of a field id (state n of the reference)
double* fieldPtr(mhy_ctx* x, int field, long& n) {
  const int c = x->cur;
  n = x->ncell;
  switch (field) {
    case MHY_CELL_DENSITY: return x->f.rho[c];
    case MHY_CELL_PRESSURE: return x->f.p[c];
    case MHY_CELL_ENERGY: return x->f.e[c];
    case MHY_CELL_PSEUDO: return x->f.q[c];
    case MHY_CELL_SOUND_SPEED: return x->f.c[c];
    case MHY_CELL_DIVU: return x->f.divu[c];
    case MHY_CELL_MASS: return x->f.M;
    case MHY_CELL_VELOCITY_X: return x->euc ? x->ef.vx[c] : nullptr;
    case MHY_CELL_VELOCITY_Y: return x->euc ? x->ef.vy[c] : nullptr;
    case MHY_CELL_OUT_VX: return x->euc ? x->ef.outvx : nullptr;
    case MHY_CELL_OUT_VY: return x->euc ? x->ef.outvy : nullptr;
    default: break;
  }
  if (field >= MHY_MAT_FRACVOL && field < MHY_MAT_FRACVOL + MHY_NBMATMAX * MHY_MAT_STRIDE) {
    const int k = (field - MHY_MAT_FRACVOL) / MHY_MAT_STRIDE, id = field - k * MHY_MAT_STRIDE;
    if (k >= x->o.nbmat) return nullptr;
    switch (id) {
      case MHY_MAT_FRACVOL: return x->f.f_k[k];
      case MHY_MAT_MASS_FRACTION: return x->f.mu_k[k];
      case MHY_MAT_DENSITY: return x->f.rho_k[c][k];
      case MHY_MAT_ENERGY: return x->f.e_k[c][k];
      case MHY_MAT_PRESSURE: return x->f.p_k[c][k];
      case MHY_MAT_PSEUDO: return x->f.q_k[c][k];
      case MHY_MAT_MASS: return x->f.M_k[k];
      case MHY_MAT_SOUND_SPEED: return x->f.c_k[c][k];
      default: return nullptr;
    }
  }
  n = x->nnode;
  switch (field) {
    case MHY_NODE_VELOCITY_X: return x->f.ux[c];
    case MHY_NODE_VELOCITY_Y: return x->f.uy[c];
    case MHY_NODE_MASS: return x->euc ? nullptr : x->f.Mp;
    case MHY_NODE_COORD_X: return x->o.mode == MODE_LAGRANGE ? x->f.X[c] : nullptr;
    case MHY_NODE_COORD_Y: return x->o.mode == MODE_LAGRANGE ? x->f.Y[c] : nullptr;
    case MHY_NODE_OUT_VX: return x->euc ? nullptr : (x->o.mode == MODE_LAGRANGE ? x->f.outvx : x->f.ux[c]);
    case MHY_NODE_OUT_VY: return x->euc ? nullptr : (x->o.mode == MODE_LAGRANGE ? x->f.outvy : x->f.uy[c]);
    default: return nullptr;
  }
}

}  // namespace

// --------------------------------------------------------------- C ABI

extern "C" {

const char* mhy_last_error(const mhy_ctx* x) { return x ? x->err.c_str() : "null context"; }

// Set-up of an Eucclhyd problem (Eucclhyd::simulate prologue,
// eucclhyd_remap/EucclhydRemap.cc:273-286): options, initial state on the
// host, device arrays (FP64), upload.  Returns MHY_OK or the failure code
// (message in x->err).
static int setupEucclhyd(mhy_ctx* x, const mhy_config& c, bool batchMember) {
  if (batchMember) return fail(x, MHY_ERR_ARG, "the Eucclhyd scheme does not run in ensembles");
  if (c.precision != 0) return fail(x, MHY_ERR_ARG, "the Eucclhyd scheme runs in FP64 only (precision 0)");
  if (c.projection_order < 1 || c.projection_order > 3)
    return fail(x, MHY_ERR_ARG, "projection order must be 1, 2 or 3");
  if (c.projection_order == 3 && c.pp == 1)
    return fail(x, MHY_ERR_ARG, "projection order 3 is not supported with the PenteBorne option");
  if (c.sans_lagrange) return fail(x, MHY_ERR_ARG, "sansLagrange is a VNR option");
  if (c.nbmat < 1 || c.nbmat > MHY_NBMATMAX) return fail(x, MHY_ERR_ARG, "1 to 3 materials");
  if (c.space_order != 1 && c.space_order != 2) return fail(x, MHY_ERR_ARG, "spaceOrder must be 1 or 2");
  for (int k = 0; k < c.nbmat; ++k)
    if (c.eos[k] != MHY_EOS_PERFECT_GAS && c.eos[k] != MHY_EOS_VOID)
      return fail(x, MHY_ERR_ARG,
                  "equation of state " + std::to_string(c.eos[k]) +
                      " is not programmed for the Eucclhyd scheme (PhaseLagrange.cc:111-135)");
  x->euc = true;
  x->mixed = false;
  x->sweepWave = true;
  x->fuseFinal = false, x->fuseLag = false, x->walkY = false;
  x->tune.on = false;
  std::string err;
  EucInit u;
  int st = buildEucInitialState(c, x->init, u, err);
  if (st != MHY_OK) return fail(x, st, err);
  if (c.avec_projection && !outerNormalsAxisExact(c, err)) return fail(x, MHY_ERR_ARG, err);
  const HostInit& h = x->init;
  x->ncell = (long)c.nx * c.ny;
  x->nnode = (long)(c.nx + 1) * (c.ny + 1);
  const long C = x->ncell, N = x->nnode;
  const int m = c.nbmat;
  Opts& o = x->o;
  std::memset(&o, 0, sizeof(o));
  o.nx = c.nx, o.ny = c.ny, o.nbmat = m;
  o.nc = 3 * m + 2 + (c.projection_conservative == 1 ? 1 : 0);
  o.mode = c.avec_projection ? MODE_FULL : MODE_LAGRANGE;
  o.euc = 1;
  o.space_order = c.space_order;
  o.conservative = (c.projection_conservative == 1) ? 1 : 0;
  o.deltat_min = c.deltat_min;
  o.order = c.projection_order;
  o.limiter = c.limiter, o.limiter_pure = c.limiter_pure;
  o.pp = c.pp, o.pp_mixte = c.pp_mixte, o.pp_complet = c.pp_complet;
  o.methode = -1;  // no nodal remap
  for (int k = 0; k < MHY_NBMATMAX; ++k) o.eos[k] = c.eos[k], o.gammap[k] = c.gammap[k];
  o.gamma = c.gamma, o.threshold = c.threshold, o.dx = c.dx, o.dy = c.dy;
  const Boundary& b = h.bc;
  o.bc_left = b.left, o.bc_right = b.right, o.bc_bottom = b.bottom, o.bc_top = b.top;
  for (int d = 0; d < 2; ++d)
    o.bcv_left[d] = b.vleft[d], o.bcv_right[d] = b.vright[d], o.bcv_bottom[d] = b.vbottom[d], o.bcv_top[d] = b.vtop[d];
  o.flux_left = b.flux_left, o.flux_right = b.flux_right, o.flux_bottom = b.flux_bottom, o.flux_top = b.flux_top;
  o.flux_any = (b.flux_left + b.flux_right + b.flux_bottom + b.flux_top) > 0;

  hipError_t e = hipSuccess;
  if ((e = hipGetDevice(&x->device)) != hipSuccess) return fail(x, MHY_ERR_DEVICE, hipGetErrorString(e));
  if (!(x->scratch = scratchAddress(x->device)))
    return fail(x, MHY_ERR_DEVICE, "hipGetSymbolAddress of the scratch array failed");
  if ((e = hipStreamCreate(&x->stream)) != hipSuccess) return fail(x, MHY_ERR_DEVICE, hipGetErrorString(e));
  Fields& f = x->f;
  EucFields& ef = x->ef;
  std::memset(&f, 0, sizeof(f));
  std::memset(&ef, 0, sizeof(ef));
#define AD(ptr, n) ((ptr) = allocD(x, (n), e))
  for (int s = 0; s < 2; ++s) {
    AD(f.rho[s], C), AD(f.e[s], C), AD(f.ux[s], N), AD(f.uy[s], N), AD(ef.vx[s], C), AD(ef.vy[s], C);
    for (int k = 0; k < m; ++k) AD(f.rho_k[s][k], C), AD(f.e_k[s][k], C);
  }
  // pressure and sound speed: one array each (the reference's m_pressure,
  // m_speed_velocity), formed at the start of every step
  AD(f.p[0], C), AD(f.c[0], C);
  f.p[1] = f.p[0], f.c[1] = f.c[0];
  for (int k = 0; k < m; ++k) {
    AD(f.p_k[0][k], C), AD(f.c_k[0][k], C);
    f.p_k[1][k] = f.p_k[0][k], f.c_k[1][k] = f.c_k[0][k];
    AD(f.f_k[k], C), AD(f.mu_k[k], C), AD(f.M_k[k], C);
    AD(ef.fsx[k], C), AD(ef.fsy[k], C), AD(ef.pmin[k], N), AD(ef.pmax[k], N);
  }
  AD(f.M, C);
  f.mixte = allocI(x, C, e);
  f.pure = allocI(x, C, e);
  for (int q = 0; q < 4; ++q) AD(ef.gv[q], C);
  for (int d = 0; d < 2; ++d) AD(ef.vmin[d], N), AD(ef.vmax[d], N);
  AD(ef.perim, C), AD(ef.outvx, C), AD(ef.outvy, C);
  if (o.mode == MODE_LAGRANGE) {
    for (int s = 0; s < 2; ++s) AD(ef.xc[s], C), AD(ef.yc[s], C), AD(ef.ve[s], C), AD(f.X[s], N), AD(f.Y[s], N);
  } else {
    AD(ef.xc[0], C), AD(ef.yc[0], C), AD(ef.ve[0], C);  // Euler cells: constant
    ef.xc[1] = ef.xc[0], ef.yc[1] = ef.yc[0], ef.ve[1] = ef.ve[0];
    for (int q = 0; q < o.nc; ++q) AD(f.UL[q], C), AD(f.U1[q], C), AD(f.U2[q], C);
    AD(f.VL, C), AD(f.XcL, C), AD(f.YcL, C);
  }
  if (e == hipSuccess) {
    void* p = nullptr;
    if ((e = hipMalloc(&p, sizeof(Scalars))) == hipSuccess) x->allocs.push_back(p);
    f.sc = static_cast<Scalars*>(p);
  }
#undef AD
  if (e != hipSuccess) return fail(x, MHY_ERR_DEVICE, std::string("device allocation: ") + hipGetErrorString(e));
  // initial state into the "n" slot 0 (setUpTimeLoopN, Init.cc:721-760)
  int ok = MHY_OK;
  auto up = [&](double* d, const std::vector<double>& v) {
    if (ok == MHY_OK) ok = upload(x, d, v);
  };
  up(f.rho[0], h.rho), up(f.e[0], h.e), up(ef.vx[0], u.vx), up(ef.vy[0], u.vy);
  for (int k = 0; k < m; ++k) {
    up(f.rho_k[0][k], h.rho_k[k]), up(f.e_k[0][k], h.e_k[k]);
    up(f.f_k[k], h.f_k[k]), up(f.mu_k[k], h.mu_k[k]);
  }
  up(ef.xc[0], u.xc), up(ef.yc[0], u.yc), up(ef.ve[0], u.ve), up(ef.perim, u.perim);
  up(ef.outvx, u.outvx), up(ef.outvy, u.outvy);
  if (o.mode == MODE_LAGRANGE) {
    std::vector<double> X(N), Y(N);
    for (long j = 0; j <= c.ny; ++j)
      for (long i = 0; i <= c.nx; ++i) {
        const long p = j * (c.nx + 1) + i;
        X[p] = c.dx * (double)i, Y[p] = c.dy * (double)j;
      }
    up(f.X[0], X), up(f.Y[0], Y);
  }
  if (ok != MHY_OK) return ok;
  if (hipMemcpy(f.mixte, h.mixte.data(), sizeof(int) * C, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(f.pure, h.pure.data(), sizeof(int) * C, hipMemcpyHostToDevice) != hipSuccess)
    return fail(x, MHY_ERR_DEVICE, "upload flags");
  Scalars s0;
  std::memset(&s0, 0, sizeof(s0));
  s0.t[0] = 0.0;
  s0.dt[0] = h.deltat_init;
  for (int q = 0; q < 2; ++q) s0.minKey[q] = dkeyPos(DBL_MAX), s0.maxKey[q] = dkeySigned(-DBL_MAX);
  if (hipMemcpy(f.sc, &s0, sizeof(s0), hipMemcpyHostToDevice) != hipSuccess)
    return fail(x, MHY_ERR_DEVICE, "upload scalars");
  x->init = HostInit();
  x->cur = 0;
  x->x_then_y_n = true;
  x->iteration = 0;
  return MHY_OK;
}

// batchMember: a problem of an ensemble, whose steps the batch launches
// (with the stand-alone Lagrange cell kernel)
static int createContext(const mhy_config* cfg, mhy_ctx** out, bool batchMember) {
  if (!cfg || !out) return MHY_ERR_ARG;
  *out = nullptr;
  mhy_ctx* x = new mhy_ctx();
  x->cfg = *cfg;
  if (const char* sk = std::getenv("MAHYCO_SWEEP_KERNEL")) x->sweepWave = (std::string(sk) != "cell");
  if (const char* ff = std::getenv("MAHYCO_FUSE_FINAL")) x->fuseFinal = (std::string(ff) != "0");
  const mhy_config& c = *cfg;
  auto bail = [&](int code) {
    // keep the context so the caller can read the message, unless it is
    // unusable; mhy_destroy releases it
    *out = x;
    return code;
  };
  // the kernels index cells and nodes with 32-bit arithmetic and address
  // every array with 32-bit byte offsets (FP64 node arrays below 4 GiB);
  // checked before any host or device allocation
  if (c.nx < 1 || c.ny < 1 || (long)(c.nx + 1) * (c.ny + 1) >= (1L << 29))
    return bail(fail(x, MHY_ERR_ARG, "mesh size out of range (at most 2^29 nodes)"));
  if (c.schema == MHY_SCHEMA_EUCCLHYD) {
    const int st = setupEucclhyd(x, c, batchMember);
    *out = x;
    return st;
  }
  if (c.schema != MHY_SCHEMA_VNR)
    return bail(fail(x, MHY_ERR_ARG, "the scheme must be VNR (2001) or Eucclhyd (2000)"));
  if (c.projection_order < 1 || c.projection_order > 3)
    return bail(fail(x, MHY_ERR_ARG, "projection order must be 1, 2 or 3"));
  if (c.projection_order == 3 && c.pp == 1)
    return bail(fail(x, MHY_ERR_ARG, "projection order 3 is not supported with the PenteBorne option"));
  if (c.precision != 0 && c.precision != 1)
    return bail(fail(x, MHY_ERR_ARG, "precision must be 0 (FP64) or 1 (mixed FP32 / FP64)"));
  x->mixed = (c.precision == 1);
  if (x->mixed) x->sweepWave = true;  // the per-cell stencil sweeps are FP64 only
  x->walkY = x->mixed && c.nbmat == 1 && c.pp != 1 && (long)sweepWalkGrid(c.nx, c.ny).x * 4 >= 4096;
  x->dualWalkY = x->mixed && (long)dualWalkGrid(c.nx, c.ny).x * 4 >= 8192;
  // FP64: the LDS-tile kernel (4096^2 point explosion: 0.42 ms against 0.45 ms
  // for the wave strip)
  x->dualTileY = !x->mixed;
  if (const char* dy = std::getenv("MAHYCO_DUAL_Y")) {
    x->dualWaveY = (std::string(dy) != "node");
    x->dualWalkY = (std::string(dy) == "walk");
    const std::string v(dy);
    x->dualTileY = (v == "tile" || v == "tile32" || v == "tile64");
    x->dualTileRows = (v == "tile32") ? 32 : (v == "tile64") ? 64 : 48;
  }
  if (const char* wy = std::getenv("MAHYCO_SWEEP_Y")) {
    if (std::string(wy) == "walk") x->walkY = true;
    if (std::string(wy) == "wave") x->walkY = false;
  }
  if (const char* g = std::getenv("MAHYCO_GRAPH")) x->useGraph = (std::string(g) != "0");
  // the fused first sweep needs the wave kernels for both sweep directions
  // (the column walker and the third-order and per-cell kernels read the
  // stored remap vector).  The PenteBorne sweeps are bound by their FP64
  // arithmetic, not by memory: the Lagrange update runs better as its own
  // streaming kernel there (2048^2 two-material case: 1.24 ms per step
  // separate, 1.27 ms fused), so it is fused with the limited-slope sweeps only.
  // (a column-walking second sweep reads the stored Lagrange geometry)
  if (const char* sg = std::getenv("MAHYCO_SWEEP_GEOM")) x->geomStored = (std::string(sg) != "recompute");
  x->fuseLag = x->sweepWave && c.projection_order != 3 && c.avec_projection == 1 && !batchMember && c.pp != 1 &&
               (!x->walkY || x->geomStored);
  if (const char* fl = std::getenv("MAHYCO_FUSE_LAGRANGE"))
    if (std::string(fl) == "0") x->fuseLag = false;
  if (const char* ly = std::getenv("MAHYCO_LAG_Y")) {
    x->fuseLagY = (std::string(ly) != "separate");
    x->tune.forcedLag = kLagNames[x->fuseLagY ? TL_FUSED : TL_SEPARATE];
    x->tune.choice = "lagrange_y=" + x->tune.forcedLag;
  }
  if (const char* sn = std::getenv("MAHYCO_SPLIT_NODE")) x->splitNode = (std::string(sn) != "0");
  // the tuner (MAHYCO_AUTOTUNE=0 disables it; a variant forced through the
  // environment is not tuned): meshes of at least 2^20 cells, where the
  // variants' speeds differ, with the wave sweeps; the first candidate of
  // each site is the default choice
  {
    bool tuneOn = (long)c.nx * c.ny >= (1L << 20) && x->sweepWave && c.projection_order != 3 &&
                  c.avec_projection == 1 && !batchMember;
    if (const char* at = std::getenv("MAHYCO_AUTOTUNE")) tuneOn = tuneOn && (std::string(at) != "0");
    auto& t = x->tune;
    if (tuneOn && !std::getenv("MAHYCO_DUAL_Y")) {
      t.cand[0][0] = x->dualTileY ? TD_TILE : (x->dualWalkY ? TD_WALK : TD_WAVE);
      int n = 1;
      // (the tile heights: 32 and 64 rows besides the default 48)
      for (int c2 : {TD_TILE, TD_WAVE, TD_WALK, TD_TILE32, TD_TILE64})
        if (c2 != t.cand[0][0]) t.cand[0][n++] = c2;
      t.ncand[0] = n;
    }
    // the walking sweep reads the stored Lagrange geometry
    const bool geom = !x->fuseLag || x->geomStored;
    if (tuneOn && geom && !std::getenv("MAHYCO_SWEEP_Y")) {
      t.cand[1][0] = x->walkY ? TS_WALK : TS_WAVE;
      t.cand[1][1] = x->walkY ? TS_WAVE : TS_WALK;
      t.ncand[1] = 2;
    }
    if (tuneOn && x->fuseLag && !std::getenv("MAHYCO_LAG_Y")) {
      t.cand[2][0] = TL_FUSED, t.cand[2][1] = TL_SEPARATE;
      t.ncand[2] = 2;
    }
    t.steps = 2 * mhy_ctx::kTuneRounds * (t.ncand[0] + t.ncand[1] + t.ncand[2]);
    t.on = t.steps > 0;
    if (t.on && (hipEventCreate(&t.ev[0]) != hipSuccess || hipEventCreate(&t.ev[1]) != hipSuccess))
      return bail(fail(x, MHY_ERR_DEVICE, "tuner events"));
  }
  if (c.sans_lagrange && !c.avec_projection)
    return bail(fail(x, MHY_ERR_ARG, "sansLagrange needs the projection"));
  std::string err;
  int st = buildInitialState(c, x->init, err);
  if (st != MHY_OK) return bail(fail(x, st, err));
  if (c.avec_projection && !outerNormalsAxisExact(c, err)) return bail(fail(x, MHY_ERR_ARG, err));

  const HostInit& h = x->init;
  x->ncell = (long)c.nx * c.ny;
  x->nnode = (long)(c.nx + 1) * (c.ny + 1);
  const long C = x->ncell, N = x->nnode;
  const int m = c.nbmat;
  Opts& o = x->o;
  std::memset(&o, 0, sizeof(o));
  o.nx = c.nx, o.ny = c.ny, o.nbmat = m, o.nc = 3 * m + 1;
  o.mode = c.sans_lagrange ? MODE_ADVECTION : (c.avec_projection ? MODE_FULL : MODE_LAGRANGE);
  o.order = c.projection_order;
  o.limiter = c.limiter, o.limiter_pure = c.limiter_pure;
  o.pp = c.pp, o.pp_mixte = c.pp_mixte, o.pp_complet = c.pp_complet;
  o.pseudo_centree = c.pseudo_centree, o.methode = c.methode_flux_masse;
  o.abort_energy = c.abort_on_energy;
  for (int k = 0; k < MHY_NBMATMAX; ++k) o.eos[k] = c.eos[k], o.gammap[k] = c.gammap[k];
  o.gamma = c.gamma, o.threshold = c.threshold, o.dx = c.dx, o.dy = c.dy;
  // computeDeltaT constants (VnrRemap.cc:16-22)
  o.cfl = c.avec_projection == 1 ? 0.05 : 0.1;
  o.cfl_euler = c.avec_projection == 1 ? 1.0 : 0.0;
  const Boundary& b = h.bc;
  o.bc_left = b.left, o.bc_right = b.right, o.bc_bottom = b.bottom, o.bc_top = b.top;
  for (int d = 0; d < 2; ++d)
    o.bcv_left[d] = b.vleft[d], o.bcv_right[d] = b.vright[d], o.bcv_bottom[d] = b.vbottom[d], o.bcv_top[d] = b.vtop[d];
  o.periodic_right = b.periodic_right;
  o.flux_left = b.flux_left, o.flux_right = b.flux_right, o.flux_bottom = b.flux_bottom, o.flux_top = b.flux_top;
  o.flux_any = (b.flux_left + b.flux_right + b.flux_bottom + b.flux_top) > 0;
  // one perfect-gas material, Lagrange + remap: the final remap stores the
  // material fields only where they differ from the means (MAHYCO_COMPACT=0
  // stores them everywhere, as for several materials)
  o.compact1 = !batchMember && m == 1 && o.mode == MODE_FULL && o.eos[0] == MHY_EOS_PERFECT_GAS;
  if (const char* cp = std::getenv("MAHYCO_COMPACT"))
    if (std::string(cp) == "0") o.compact1 = 0;

  hipError_t e = hipSuccess;
  if ((e = hipGetDevice(&x->device)) != hipSuccess) return bail(fail(x, MHY_ERR_DEVICE, hipGetErrorString(e)));
  if (!(x->scratch = scratchAddress(x->device)))
    return bail(fail(x, MHY_ERR_DEVICE, "hipGetSymbolAddress of the scratch array failed"));
  if ((e = hipStreamCreate(&x->stream)) != hipSuccess) return bail(fail(x, MHY_ERR_DEVICE, hipGetErrorString(e)));
  Fields& f = x->f;
  std::memset(&f, 0, sizeof(f));
#define AD(ptr, n) ((ptr) = allocD(x, (n), e))
#define AR(ptr, n) ((ptr) = allocR(x, (n), e))
  for (int s = 0; s < 2; ++s) {
    AR(f.rho[s], C), AR(f.p[s], C), AR(f.q[s], C), AR(f.c[s], C), AR(f.divu[s], C), AR(f.e[s], C);
    for (int k = 0; k < m; ++k) {
      AR(f.rho_k[s][k], C), AR(f.p_k[s][k], C), AR(f.q_k[s][k], C), AR(f.c_k[s][k], C), AR(f.e_k[s][k], C);
    }
    AR(f.ux[s], N), AR(f.uy[s], N);
  }
  if (o.mode == MODE_LAGRANGE) {
    for (int s = 0; s < 2; ++s) AD(f.X[s], N), AD(f.Y[s], N), AD(f.vol[s], C);
    AR(f.outvx, N), AR(f.outvy, N);
  } else {
    AD(f.vol[0], C);
    f.vol[1] = f.vol[0];  // Euler sub-volumes: constant
  }
  for (int k = 0; k < m; ++k) AD(f.f_k[k], C), AR(f.mu_k[k], C), AD(f.M_k[k], C);
  AD(f.M, C);
  f.mixte = allocI(x, C, e);
  f.pure = allocI(x, C, e);
  if (o.compact1 && e == hipSuccess) {
    void* p = nullptr;
    if ((e = hipMalloc(&p, C)) == hipSuccess) x->allocs.push_back(p);
    f.envx = static_cast<unsigned char*>(p);
  }
  AR(f.Mp, N);
  if (o.periodic_right) {
    AR(f.rho0, C), AR(f.e0, C);
    for (int k = 0; k < m; ++k) AR(f.rho_k0[k], C), AR(f.e_k0[k], C);
  }
  if (o.mode != MODE_LAGRANGE) {
    // U_L, V_L and the centroid: the fused first sweep keeps U_L in
    // registers, but a step whose first sweep runs along y may run the
    // separate Lagrange kernel (the tuner's choice)
    for (int q = 0; q < o.nc; ++q) AD(f.UL[q], C), AD(f.U1[q], C), AD(f.U2[q], C);
    AD(f.VL, C), AD(f.XcL, C), AD(f.YcL, C);
    if (o.methode == 1) AD(f.rL, C);
    if (o.pp == 1) AD(f.qL, C);
    x->nface = std::max((long)(c.nx + 1) * c.ny, (long)c.nx * (c.ny + 1));
    if (o.methode == 0 || o.methode == 1)
      for (int k = 0; k < m; ++k) AD(f.G[k], x->nface);
    if (o.methode == 2)
      for (int k = 0; k < m; ++k) AD(f.DF[k], C);
    for (int d = 0; d < 3; ++d) AD(f.UD1[d], N);
    AD(f.XL2, N);
  }
  if (e == hipSuccess) {
    void* p = nullptr;
    if ((e = hipMalloc(&p, sizeof(Scalars))) == hipSuccess) x->allocs.push_back(p);
    f.sc = static_cast<Scalars*>(p);
  }
#undef AR
#undef AD
  if (e != hipSuccess) return bail(fail(x, MHY_ERR_DEVICE, std::string("device allocation: ") + hipGetErrorString(e)));

  // initial state into the "n" slot 0 (setUpTimeLoopN, VnrRemap.cc:68-87)
  int u = MHY_OK;
  auto up = [&](double* d, const std::vector<double>& s) {
    if (u == MHY_OK) u = upload(x, d, s);
  };
  auto upR = [&](double* d, const std::vector<double>& s) {
    if (u == MHY_OK) u = uploadR(x, d, s);
  };
  upR(f.rho[0], h.rho), upR(f.p[0], h.p), upR(f.c[0], h.c), upR(f.e[0], h.e);
  for (int k = 0; k < m; ++k) {
    upR(f.rho_k[0][k], h.rho_k[k]), upR(f.p_k[0][k], h.p_k[k]), upR(f.c_k[0][k], h.c_k[k]);
    upR(f.e_k[0][k], h.e_k[k]);
    up(f.f_k[k], h.f_k[k]), upR(f.mu_k[k], h.mu_k[k]), up(f.M_k[k], h.M_k[k]);
  }
  up(f.M, h.M);
  up(f.vol[0], h.vol);
  upR(f.ux[0], h.ux), upR(f.uy[0], h.uy);
  if (o.mode == MODE_LAGRANGE) {
    std::vector<double> X(N), Y(N);
    for (long j = 0; j <= c.ny; ++j)
      for (long i = 0; i <= c.nx; ++i) {
        const long p = j * (c.nx + 1) + i;
        X[p] = c.dx * (double)i, Y[p] = c.dy * (double)j;
      }
    up(f.X[0], X), up(f.Y[0], Y);
    upR(f.outvx, h.ux), upR(f.outvy, h.uy);
  }
  if (o.periodic_right) {
    upR(f.rho0, h.rho), upR(f.e0, h.e);
    for (int k = 0; k < m; ++k) upR(f.rho_k0[k], h.rho_k[k]), upR(f.e_k0[k], h.e_k[k]);
  }
  if (u != MHY_OK) return bail(u);
  if (hipMemcpy(f.mixte, h.mixte.data(), sizeof(int) * C, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(f.pure, h.pure.data(), sizeof(int) * C, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(x, MHY_ERR_DEVICE, "upload flags"));
  // the initial material fields are stored: every cell reads them
  if (o.compact1 && hipMemset(f.envx, 1, C) != hipSuccess) return bail(fail(x, MHY_ERR_DEVICE, "material flags"));
  Scalars s0;
  std::memset(&s0, 0, sizeof(s0));
  s0.t[0] = 0.0;
  s0.dt[0] = h.deltat_init;
  for (int q = 0; q < 2; ++q) s0.minKey[q] = dkeyPos(DBL_MAX), s0.maxKey[q] = dkeySigned(-DBL_MAX);
  if (hipMemcpy(f.sc, &s0, sizeof(s0), hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(x, MHY_ERR_DEVICE, "upload scalars"));
  // computeNodeMass of the prologue (VnrRemap.cc:341)
  {
    KP kp = makeKP(x);
    launchR(x, k_node_mass<double>, k_node_mass<float>, gridFor(N), dim3(kBlock), kp);
    if (hipStreamSynchronize(x->stream) != hipSuccess) return bail(fail(x, MHY_ERR_DEVICE, "node mass"));
  }
  x->init = HostInit();  // the host copy of the initial state is no longer needed
  x->cur = 0;
  x->x_then_y_n = true;
  x->iteration = 0;
  *out = x;
  return MHY_OK;
}

int mhy_create(const mhy_config* cfg, mhy_ctx** out) { return createContext(cfg, out, false); }

int mhy_destroy(mhy_ctx* x) {
  if (!x) return MHY_OK;
  if (x->stream) (void)hipStreamSynchronize(x->stream);
  for (auto& r : x->pending) (void)hipEventDestroy(r.a), (void)hipEventDestroy(r.b);
  for (auto ev : x->eventPool) (void)hipEventDestroy(ev);
  for (auto& g : x->graph)
    if (g) (void)hipGraphExecDestroy(g);
  for (auto ev : x->tune.ev)
    if (ev) (void)hipEventDestroy(ev);
  for (void* p : x->allocs) (void)hipFree(p);
  if (x->stream) (void)hipStreamDestroy(x->stream);
  delete x->writer;
  delete x;
  return MHY_OK;
}

int mhy_step(mhy_ctx* x, int sync) {
  if (!x) return MHY_ERR_ARG;
  if (x->status != MHY_OK) return x->status;
  if (x->finished) return fail(x, MHY_ERR_FINISHED, "time loop already ended");
  int st = runIteration(x);
  if (st != MHY_OK) return st;
  endIteration(x);
  if (sync) {
    Scalars s;
    if ((st = readScalars(x, s)) != MHY_OK) return st;
    if (s.err) return fail(x, s.err, "kernel status " + std::to_string(s.err));
  }
  return MHY_OK;
}

// Captures the step that starts with the current parity and sweep order.
// Capturing records the launches without running them, so the host-side
// step state (iteration count, fused-final flag) is put back afterwards.
int captureStep(mhy_ctx* x) {
  const long it = x->iteration;
  hipGraph_t g = nullptr;
  HIPCHK(x, hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal));
  const int st = runIteration(x);
  const hipError_t e = hipStreamEndCapture(x->stream, &g);
  x->iteration = it;
  x->finalDone = false;
  if (st != MHY_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) return fail(x, MHY_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(&x->graph[2 * x->cur + (x->x_then_y_n ? 1 : 0)], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) return fail(x, MHY_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(ei));
  return MHY_OK;
}

// One step (without the n/n+1 swap): a replay of the captured step when
// graphs is set, else eager launches.
int stepOnce(mhy_ctx* x, bool graphs) {
  // no graph is captured before the tuner has fixed the kernel variants
  if (!graphs || x->tune.on) return runIteration(x);
  hipGraphExec_t& g = x->graph[2 * x->cur + (x->x_then_y_n ? 1 : 0)];
  if (!g) {
    const int st = captureStep(x);
    if (st != MHY_OK) return st;
  }
  HIPCHK(x, hipGraphLaunch(g, x->stream));
  x->iteration++;
  return MHY_OK;
}

int mhy_run(mhy_ctx* x, int nsteps) {
  if (!x) return MHY_ERR_ARG;
  if (x->status != MHY_OK) return x->status;
  if (x->finished) return fail(x, MHY_ERR_FINISHED, "time loop already ended");
  // graphs need the launches of a step to depend only on the parity; kernel
  // events (timing) are recorded on the eager path
  const bool graphs = x->useGraph && !x->timing && nsteps >= 2;
  for (int s = 0; s < nsteps; ++s) {
    const int st = stepOnce(x, graphs);
    if (st != MHY_OK) return st;
    endIteration(x);
  }
  Scalars s;
  int st = readScalars(x, s);
  if (st != MHY_OK) return st;
  if (s.err) return fail(x, s.err, "kernel status " + std::to_string(s.err));
  return MHY_OK;
}

int mhy_compute_dt(mhy_ctx* x) { return x ? phaseDt(x) : MHY_ERR_ARG; }
int mhy_lagrange(mhy_ctx* x) { return x ? phaseLagrange(x) : MHY_ERR_ARG; }
// the remap preparation is fused into the cell kernel of mhy_lagrange
int mhy_prepare_remap(mhy_ctx* x) { return x ? MHY_OK : MHY_ERR_ARG; }
int mhy_remap_sweep(mhy_ctx* x, int sweep) {
  if (!x || (sweep != 1 && sweep != 2)) return MHY_ERR_ARG;
  if (x->o.mode == MODE_LAGRANGE) return MHY_OK;
  return phaseSweep(x, sweep);
}
int mhy_remap_stage(mhy_ctx* x, int sweep, int stage) {
  if (!x || (sweep != 1 && sweep != 2) || (stage != MHY_STAGE_CELLS && stage != MHY_STAGE_NODES))
    return MHY_ERR_ARG;
  if (x->o.mode == MODE_LAGRANGE) return MHY_OK;
  return phaseSweepStage(x, sweep, stage);
}
int mhy_output_open(mhy_ctx* x, const char* output_dir) {
  if (!x) return MHY_ERR_ARG;
  delete x->writer;
  x->writer = new OutputWriter(output_dir ? output_dir : "none");
  return MHY_OK;
}
int mhy_output_write(mhy_ctx* x, int index, double t) {
  if (!x || !x->writer) return MHY_ERR_ARG;
  if (x->status != MHY_OK) return x->status;
  return x->writer->write(x, index, t);
}
int mhy_remap_final(mhy_ctx* x) {
  if (!x) return MHY_ERR_ARG;
  if (x->o.mode == MODE_LAGRANGE) return MHY_OK;
  return phaseFinal(x);
}
int mhy_end_iteration(mhy_ctx* x) {
  if (!x) return MHY_ERR_ARG;
  x->iteration++;
  endIteration(x);
  return MHY_OK;
}

int mhy_synchronize(mhy_ctx* x) {
  if (!x) return MHY_ERR_ARG;
  HIPCHK(x, hipStreamSynchronize(x->stream));
  collectTimes(x);
  return MHY_OK;
}

int mhy_get_time_state(mhy_ctx* x, mhy_time_state* ts) {
  if (!x || !ts) return MHY_ERR_ARG;
  Scalars s;
  int st = readScalars(x, s);
  if (st != MHY_OK) return st;
  ts->t_n = s.t[x->cur], ts->t_nplus1 = s.t[1 - x->cur];
  ts->deltat_n = s.dt[x->cur], ts->deltat_nplus1 = s.dt[1 - x->cur];
  ts->iteration = x->iteration;
  ts->x_then_y_n = x->x_then_y_n ? 1 : 0;
  ts->finished = x->finished ? 1 : 0;
  ts->status = s.err ? s.err : x->status;
  return MHY_OK;
}

int mhy_get_mesh_size(const mhy_ctx* x, int* nx, int* ny) {
  if (!x) return MHY_ERR_ARG;
  if (nx) *nx = x->o.nx;
  if (ny) *ny = x->o.ny;
  return MHY_OK;
}

int mhy_get_field(mhy_ctx* x, int field, double* out, long long n) {
  if (!x || !out) return MHY_ERR_ARG;
  if (isMaterialField(field)) {
    const int st = expandMaterialFields(x);
    if (st != MHY_OK) return st;
  }
  HIPCHK(x, hipStreamSynchronize(x->stream));
  if (field == MHY_CELL_MIXTE || field == MHY_CELL_PURE) {
    if (n != x->ncell) return fail(x, MHY_ERR_ARG, "size");
    std::vector<int> v(x->ncell);
    HIPCHK(x, hipMemcpy(v.data(), field == MHY_CELL_MIXTE ? x->f.mixte : x->f.pure, sizeof(int) * x->ncell,
                        hipMemcpyDeviceToHost));
    for (long c = 0; c < x->ncell; ++c) out[c] = v[c];
    return MHY_OK;
  }
  if (field == MHY_NODE_MASS) {
    if (x->euc) return fail(x, MHY_ERR_ARG, "the Eucclhyd scheme has no node mass");
    KP kp = makeKP(x);
    launchR(x, k_node_mass<double>, k_node_mass<float>, gridFor(x->nnode), dim3(kBlock), kp);
    HIPCHK(x, hipStreamSynchronize(x->stream));
  }
  if ((field == MHY_NODE_COORD_X || field == MHY_NODE_COORD_Y) && x->o.mode != MODE_LAGRANGE) {
    if (n != x->nnode) return fail(x, MHY_ERR_ARG, "size");
    for (long j = 0; j <= x->o.ny; ++j)
      for (long i = 0; i <= x->o.nx; ++i)
        out[j * (x->o.nx + 1) + i] = (field == MHY_NODE_COORD_X) ? x->o.dx * (double)i : x->o.dy * (double)j;
    return MHY_OK;
  }
  long len = 0;
  double* p = fieldPtr(x, field, len);
  if (!p) return fail(x, MHY_ERR_ARG, "unknown field " + std::to_string(field));
  if (n != len) return fail(x, MHY_ERR_ARG, "size");
  if (x->mixed && isBulkField(field)) {
    std::vector<float> h(len);
    HIPCHK(x, hipMemcpy(h.data(), p, sizeof(float) * len, hipMemcpyDeviceToHost));
    for (long i = 0; i < len; ++i) out[i] = h[i];
    return MHY_OK;
  }
  HIPCHK(x, hipMemcpy(out, p, sizeof(double) * len, hipMemcpyDeviceToHost));
  return MHY_OK;
}

int mhy_set_field(mhy_ctx* x, int field, const double* in, long long n) {
  if (!x || !in) return MHY_ERR_ARG;
  long len = 0;
  double* p = fieldPtr(x, field, len);
  if (!p || field == MHY_NODE_MASS) return fail(x, MHY_ERR_ARG, "field not settable");
  if (n != len) return fail(x, MHY_ERR_ARG, "size");
  // a compact context stores every cell's material fields from here on, so
  // that setting a mean field leaves the material fields as they were
  const int st = storeMaterialFields(x);
  if (st != MHY_OK) return st;
  HIPCHK(x, hipStreamSynchronize(x->stream));
  if (x->mixed && isBulkField(field)) {
    std::vector<float> h(in, in + len);
    HIPCHK(x, hipMemcpy(p, h.data(), sizeof(float) * len, hipMemcpyHostToDevice));
    return MHY_OK;
  }
  HIPCHK(x, hipMemcpy(p, in, sizeof(double) * len, hipMemcpyHostToDevice));
  return MHY_OK;
}

}  // extern "C"

namespace {

// Neumaier sum of the block partials (sum and compensation of both passes)
// of each quantity: the final step of both conservation entry points.
void consFinish(const double* h, unsigned nb, int nq, double* sums) {
  for (int q = 0; q < nq; ++q) {
    double t = 0.0, cc = 0.0;
    for (unsigned bl = 0; bl < nb; ++bl)
      for (int part = 0; part < 4; ++part) {
        const double v = h[(bl * nq + q) * 4 + part];
        const double y = t + v;
        if (std::fabs(t) >= std::fabs(v)) cc += (t - y) + v;
        else cc += (v - y) + t;
        t = y;
      }
    sums[q] = t + cc;
  }
}

unsigned consBlocks(const mhy_ctx* x) {
  return (unsigned)std::min<long>(kConsBlocks, (x->ncell + kBlock - 1) / kBlock);
}

}  // namespace

extern "C" {

int mhy_conservation(mhy_ctx* x, double* sums) {
  if (!x || !sums) return MHY_ERR_ARG;
  if (x->euc) return fail(x, MHY_ERR_ARG, "conservation sums are formed for the VNR scheme");
  const int m = x->o.nbmat, nq = 2 + m;
  const unsigned nb = consBlocks(x);
  double* d = nullptr;
  HIPCHK(x, hipMalloc(&d, sizeof(double) * nb * nq * 4));
  KP kp = makeKP(x);
  launchR(x, k_conservation<double>, k_conservation<float>, dim3(nb), dim3(kBlock), kp, d);
  std::vector<double> h(nb * nq * 4);
  hipError_t e1 = hipMemcpyAsync(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost, x->stream);
  hipError_t e2 = hipStreamSynchronize(x->stream);
  (void)hipFree(d);
  if (e1 != hipSuccess || e2 != hipSuccess) return fail(x, MHY_ERR_DEVICE, "conservation");
  consFinish(h.data(), nb, nq, sums);
  return MHY_OK;
}

// The same sums with the block partials in this device's copy of the
// scratch array instead of a buffer allocated per call.  Every slot the
// finish reads is written by the kernel of the same call, so nothing is
// cleared.  Calls on the same GPU share the array: calls that overlap (from
// several host threads or streams) race.
int mhy_conservation_noalloc(mhy_ctx* x, double* sums) {
  if (!x || !sums) return MHY_ERR_ARG;
  if (x->euc) return fail(x, MHY_ERR_ARG, "conservation sums are formed for the VNR scheme");
  const int m = x->o.nbmat, nq = 2 + m;
  const unsigned nb = consBlocks(x);
  const size_t need = (size_t)nb * nq * 4;
  if (need > (size_t)kConsScratch)
    return fail(x, MHY_ERR_ARG, "conservation partials (" + std::to_string(need) + " doubles) exceed the scratch "
                                "array (" + std::to_string(kConsScratch) + ")");
  if (!x->scratch) return fail(x, MHY_ERR_DEVICE, "scratch array not set up on this device");
  KP kp = makeKP(x);
  launchR(x, k_conservation<double>, k_conservation<float>, dim3(nb), dim3(kBlock), kp, x->scratch);
  double h[kConsScratch];
  hipError_t e1 = hipMemcpyAsync(h, x->scratch, sizeof(double) * need, hipMemcpyDeviceToHost, x->stream);
  hipError_t e2 = hipStreamSynchronize(x->stream);
  if (e1 != hipSuccess || e2 != hipSuccess) return fail(x, MHY_ERR_DEVICE, "conservation");
  consFinish(h, nb, nq, sums);
  return MHY_OK;
}

const char* mhy_tuned_kernels(const mhy_ctx* x) { return x ? x->tune.choice.c_str() : ""; }

int mhy_kernel_count(const mhy_ctx*) { return K_COUNT; }
const char* mhy_kernel_name(const mhy_ctx*, int k) { return (k >= 0 && k < K_COUNT) ? kKernelNames[k] : ""; }

// Algorithmic bytes of one launch: every input word the kernel must read
// from HBM once and every output word it must write once (neighbour reuse
// through the caches counted once).  Bulk fields count 4 bytes in the mixed
// form, everything else 8 (flags 4).
int mhy_kernel_bytes(const mhy_ctx* x, int k, double* bytes) {
  if (!x || !bytes || k < 0 || k >= K_COUNT) return MHY_ERR_ARG;
  const double C = (double)x->ncell, N = (double)x->nnode, F = (double)x->nface;
  const int m = x->o.nbmat, nc = x->o.nc;
  const bool pp = x->o.pp == 1;
  const int meth = x->o.methode;
  const double r = x->mixed ? 4.0 : 8.0, d = 8.0, i4 = 4.0;
  // cell part of the final remap: U2 in registers or read; writes rho, e, q,
  // p, c, mu_k, rho_k, e_k, p_k, c_k (bulk) and M, f_k, M_k, flags
  const double finalWrites = C * r * (5 + 5 * m) + C * d * (1 + 2 * m) + C * i4 * 2;
  double w = 0;
  if (x->euc) {  // Eucclhyd kernels (FP64)
    const bool o2 = x->o.space_order == 2, lag = x->o.mode == MODE_LAGRANGE;
    // what the corner extrapolation reads of a cell: p_k, f_k, rho_k, the
    // force sums, velocity, gradient, centroid, p, rho, c, volume
    const double cellX = C * d * (5 * m + 12);
    const double ranges = o2 ? N * d * (2 * m + 4) : 0.0;
    switch (k) {
      case K_DT:  // rho, e_k, rho_k, f_k, mu_k, c, volume, perimeter, velocity (+ node u) ;
                  // p, c, p_k, c_k, gradient, M, M_k
        w = C * d * (3 + 3 * m + 2 + 2) + (o2 ? N * 2 * d : 0.0) + C * d * (2 + 2 * m + (o2 ? 4 : 0) + 1 + m);
        break;
      case K_NODE: w = cellX + N * d * 2 + ranges + (lag ? N * d * 4 : 0.0); break;
      case K_CELL:  // + M, M_k, mu_k, e_k and the node velocities and ranges ; U_L or state n+1, force sums
        w = cellX + C * d * (1 + 3 * m) + N * d * 2 + ranges + C * d * 2 * m +
            (lag ? C * d * (2 + 2 + 2 * m + 2 + 3 + 1) : C * d * (nc + 3));
        break;
      case K_SWEEP_X:
      case K_SWEEP_Y:  // U_in, V_L and centroid, node velocities (+ flags) ; U_out
        w = C * d * nc + C * d * 2 + N * 2 * d + C * i4 * 2 * (pp && x->o.pp_mixte) + C * d * nc;
        break;
      case K_FINAL:  // U_2, volume ; rho, e, velocity, f_k, mu_k, rho_k, e_k, outputs, flags
        w = C * d * (nc + 1) + C * d * (2 + 2 + 4 * m + 2) + C * i4 * 2;
        break;
      default: w = 0.0; break;
    }
    *bytes = w;
    return MHY_OK;
  }
  switch (k) {
    case K_DT: w = C * (d + 2 * r) + N * 2 * r; break;  // vol ; c, p ; u
    case K_NODE: w = N * 2 * r + C * (2 * r + d) + N * 3 * r; break;  // u_n ; p, q ; M ; u_n+1, Mp
    case K_CELL:
      // u_n+1 ; f, M_k, vol (FP64), mu, rho_k, e_k, p_k, q_k + rho, divu, c ;
      // writes UL, VL, XcL, YcL (+ rL, qL) (FP64), divu, q_k
      w = N * 2 * r + C * ((2 * m + 1) * d + (5 * m + 3) * r) + C * d * (nc + 3 + (meth == 1) + pp) + C * r * (1 + m);
      break;
    case K_SWEEP_X:
    case K_SWEEP_Y: {
      // U_in (+ qL, flags) ; both node velocity components (the wave sweeps
      // form V_L and the centroid from the node positions; the other kernels
      // read them) ; writes U_out, face fluxes (or PB)
      const double geo = x->sweepWave && x->o.order != 3 ? 0.0 : C * d * 2;
      const double in = C * d * nc + geo + C * d * pp + C * i4 * 2 * (pp && x->o.pp_mixte) + N * 2 * r;
      const double fl = (meth == 2 ? C * m : F * m) * d;
      const double s1 = in + C * nc * d + fl;
      // sweep 2 with the final remap fused: no U2 store, the final outputs
      // instead; X and Y alternate between sweep 1 and sweep 2, unless
      // sweep 1 is the fused Lagrange sweep (its own kernel id)
      const double s2 = x->fuseFinal ? in + finalWrites + fl : s1;
      // a sweep direction whose first sweeps are the fused Lagrange sweep
      // (their own kernel id) runs only second sweeps under this id
      const bool lagOwn = x->fuseLag && (k == K_SWEEP_X || x->fuseLagY);
      w = lagOwn ? s2 : 0.5 * (s1 + s2);
      break;
    }
    case K_LAGSWEEP_X:
    case K_LAGSWEEP_Y: {
      // the Lagrange cell update's inputs (u_n+1 ; f, M_k, vol (FP64), mu,
      // rho_k, e_k, p_k, q_k + rho, divu, c) and outputs that outlive the
      // step (divu, q_k ; rL, qL), then sweep 1's outputs: U_1, face fluxes
      const double lagIn = N * 2 * r + C * ((2 * m + 1) * d + (5 * m + 3) * r);
      const double lagOut = C * r * (1 + m) + C * d * ((meth == 1) + pp);
      const double fl = (meth == 2 ? C * m : F * m) * d;
      w = lagIn + lagOut + C * nc * d + fl;
      break;
    }
    case K_DUAL_X:
    case K_DUAL_Y:
      // face fluxes (or PB cell fluxes) + sweep 1: u, Mp ; UD1, XL2
      // or sweep 2: UD1, XL2 ; u (average of the two)
      w = (meth == 2 ? C * m : F * m) * d + 0.5 * (N * 3 * r + N * 4 * d) + 0.5 * (N * 4 * d + N * 2 * r);
      break;
    case K_FINAL:
      // U2 (+ void EOS inputs) ; final outputs
      w = C * nc * d + finalWrites;
      break;
  }
  *bytes = w;
  return MHY_OK;
}

int mhy_enable_kernel_timing(mhy_ctx* x, int on) {
  if (!x) return MHY_ERR_ARG;
  (void)hipStreamSynchronize(x->stream);
  collectTimes(x);
  x->timing = (on != 0);
  for (int k = 0; k < K_COUNT; ++k) x->ms[k] = 0, x->launches[k] = 0;
  return MHY_OK;
}

int mhy_set_kernel_timing_mask(mhy_ctx* x, unsigned mask) {
  if (!x) return MHY_ERR_ARG;
  x->timingMask = mask;
  return MHY_OK;
}

int mhy_kernel_time(mhy_ctx* x, int k, double* avg_ms, int* launches) {
  if (!x || k < 0 || k >= K_COUNT) return MHY_ERR_ARG;
  (void)hipStreamSynchronize(x->stream);
  collectTimes(x);
  if (avg_ms) *avg_ms = x->launches[k] ? x->ms[k] / x->launches[k] : 0.0;
  if (launches) *launches = x->launches[k];
  return MHY_OK;
}

// The reference time loop with output (VnrRemap.cc:127-253): the last
// iteration keeps its n+1 buffers unswapped, and dumps happen after
// computeTime with the state of the step start.
int mhy_simulate(mhy_ctx* x, const char* output_dir, int verbose) {
  if (!x) return MHY_ERR_ARG;
  if (x->status != MHY_OK) return x->status;
  if (x->finished) return fail(x, MHY_ERR_FINISHED, "time loop already ended");
  std::string od = output_dir ? output_dir : "none";
  delete x->writer;
  x->writer = new OutputWriter(od);
  double lastDump = 0.0;
  int nbCalls = 0;
  x->iteration = 0;
  bool cont = true;
  std::vector<double> buf;
  auto dump = [&](double tn) -> int {
    nbCalls++;
    if (x->writer->disabled() || !(tn >= lastDump + x->cfg.output_time || tn == 0.)) return MHY_OK;
    if (x->euc) {
      // the Eucclhyd dump prints the pressure of state n, which the step
      // forms at its start (dumpVariables after computeEOS,
      // EucclhydRemap.cc:99-106); the EOS kernel is idempotent
      KP kp = makeKP(x);
      static const KernelFn ke[3] = {k_euc_eos<1>, k_euc_eos<2>, k_euc_eos<3>};
      hipLaunchKernelGGL(ke[x->o.nbmat - 1], gridFor(x->ncell), dim3(kBlock), 0, x->stream, kp);
      HIPCHK(x, hipStreamSynchronize(x->stream));
    }
    int st = x->writer->write(x, nbCalls, tn);
    lastDump = tn;
    return st;
  };
  // the dump at t_n only reads fields that the dt phase leaves unchanged, so
  // it comes before the step, which then runs as one graph replay
  const bool graphs = x->useGraph && !x->timing;
  // the time scalars read after a step are those the next step starts from
  // (nothing runs on the device in between), so one read per step serves both
  Scalars s;
  {
    const int st = readScalars(x, s);
    if (st != MHY_OK) return st;
  }
  do {
    int st;
    const double tn = s.t[x->cur];
    if ((st = dump(tn)) != MHY_OK) return st;
    if ((st = stepOnce(x, graphs)) != MHY_OK) return st;
    if ((st = readScalars(x, s)) != MHY_OK) return st;
    if (s.err) return fail(x, s.err, "kernel status " + std::to_string(s.err));
    const double t1 = s.t[1 - x->cur];
    cont = (x->iteration + 1 < x->cfg.max_time_iterations && t1 < x->cfg.final_time);
    if (verbose)
      std::printf("[%5d] t = %.8e DT = %.8e\n", x->iteration, tn, s.dt[1 - x->cur]);
    if (cont) endIteration(x);
  } while (cont);
  if (!x->euc) {  // Vnr::simulate dumps once more; Eucclhyd::simulate does not
    Scalars s;
    int st = readScalars(x, s);
    if (st != MHY_OK) return st;
    if ((st = dump(s.t[x->cur])) != MHY_OK) return st;
  }
  x->finished = true;
  return MHY_OK;
}

}  // extern "C"

// Field access for the output writer.
namespace mhy {
int ctxField(mhy_ctx* x, int field, std::vector<double>& out) {
  long n = (field >= MHY_NODE_VELOCITY_X) ? x->nnode : x->ncell;
  out.resize(n);
  return mhy_get_field(x, field, out.data(), n);
}
int ctxNbmat(const mhy_ctx* x) { return x->o.nbmat; }
int ctxProjection(const mhy_ctx* x) { return x->cfg.avec_projection; }
int ctxEuc(const mhy_ctx* x) { return x->euc ? 1 : 0; }
void ctxMesh(const mhy_ctx* x, long& nx, long& ny) { nx = x->o.nx, ny = x->o.ny; }
}  // namespace mhy

// ------------------------------------------------------------------ batches
// An ensemble of independent problems of any sizes, stepped together: every
// kernel of a step is one launch over the virtual blocks of all problems
// (BatchLaunch), each problem keeping its own arrays, dt and time.  The
// number of workgroups of each launch is chosen by timing candidates over the
// first steps (results do not depend on it); the step is then replayed from
// two captured HIP graphs (even and odd steps).
namespace {
using namespace mhy;

enum BatchKernel { BK_DT = 0, BK_NODE, BK_CELL, BK_SWX, BK_SWY, BK_DUX, BK_DUY, BK_COUNT };
constexpr int kVariants = 16;  // cur (1 bit), sweep-1 (1), xdir (1), fuseFinal (1)
constexpr unsigned kCand[] = {0u, 1024u, 2048u, 4096u, 8192u};  // 0: one workgroup per virtual block
constexpr int kNCand = sizeof(kCand) / sizeof(kCand[0]);
constexpr int kTuneSteps = 2 * kNCand;  // each candidate over two steps
template <class R, bool ONE>
void (*sweepBatchKernel(bool xd, int m, bool pp))(BatchLaunchT<R>) {
  using Fn = void (*)(BatchLaunchT<R>);
  static const Fn t[2][3][2] = {
      {{k_sweep_wave_batch<R, 1, false, false, ONE>, k_sweep_wave_batch<R, 1, true, false, ONE>},
       {k_sweep_wave_batch<R, 2, false, false, ONE>, k_sweep_wave_batch<R, 2, true, false, ONE>},
       {k_sweep_wave_batch<R, 3, false, false, ONE>, k_sweep_wave_batch<R, 3, true, false, ONE>}},
      {{k_sweep_wave_batch<R, 1, false, true, ONE>, k_sweep_wave_batch<R, 1, true, true, ONE>},
       {k_sweep_wave_batch<R, 2, false, true, ONE>, k_sweep_wave_batch<R, 2, true, true, ONE>},
       {k_sweep_wave_batch<R, 3, false, true, ONE>, k_sweep_wave_batch<R, 3, true, true, ONE>}}};
  return t[xd ? 1 : 0][m - 1][pp ? 1 : 0];
}
template <class R, bool ONE>
void (*o3BatchKernel(bool xd, int m))(BatchLaunchT<R>) {
  using Fn = void (*)(BatchLaunchT<R>);
  static const Fn t[2][3] = {
      {k_sweep_o3_batch<R, 1, false, ONE>, k_sweep_o3_batch<R, 2, false, ONE>, k_sweep_o3_batch<R, 3, false, ONE>},
      {k_sweep_o3_batch<R, 1, true, ONE>, k_sweep_o3_batch<R, 2, true, ONE>, k_sweep_o3_batch<R, 3, true, ONE>}};
  return t[xd ? 1 : 0][m - 1];
}
template <class R, bool ONE>
void (*walkBatchKernel(int m, bool pp))(BatchLaunchT<R>) {
  using Fn = void (*)(BatchLaunchT<R>);
  static const Fn t[3][2] = {{k_sweep_walk_batch<R, 1, false, ONE>, k_sweep_walk_batch<R, 1, true, ONE>},
                             {k_sweep_walk_batch<R, 2, false, ONE>, k_sweep_walk_batch<R, 2, true, ONE>},
                             {k_sweep_walk_batch<R, 3, false, ONE>, k_sweep_walk_batch<R, 3, true, ONE>}};
  return t[m - 1][pp ? 1 : 0];
}
}  // namespace

struct mhy_batch {
  std::vector<mhy_ctx*> ctx;
  hipStream_t stream = nullptr;
  int m = 0;
  bool pp = false;
  bool mixed = false;  // every problem in the mixed FP32 / FP64 form
  bool dualWaveY = true;  // nodal remap along y with the wave kernel
  bool walkY = false;     // cell sweeps along y with the column-walking kernel
  bool o3 = false;        // third-order remap (LDS-tile cell sweeps)
  void* dTable = nullptr;  // KPT<double> or KPT<float> entries
  unsigned* dStart = nullptr;  // BK_COUNT x (n+1)
  unsigned* dGx = nullptr;     // BK_COUNT x n
  unsigned* dOwner = nullptr;  // problem of each virtual block, kernel after kernel
  size_t ownerOff[BK_COUNT] = {};
  unsigned total[BK_COUNT] = {};
  unsigned blocks[BK_COUNT] = {};  // workgroups per launch (0: total)
  int cur = 0, cur0 = 0;
  bool xty = true, xty0 = true;
  long iteration = 0;
  long long cells = 0;
  bool tune = true, useGraph = true;
  int tuneStep = 0;
  float tuneMs[BK_COUNT][kNCand] = {};
  hipEvent_t ev[2 * BK_COUNT] = {};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  std::string err;
};

namespace {

int batchFail(mhy_batch* b, int code, const std::string& msg) {
  b->err = msg;
  return code;
}

// one step of every problem with bulk storage type R; timing != 0 records
// events around each launch
template <class R>
int batchEnqueueStepR(mhy_batch* b, int cur, bool xty, bool timing) {
  const int n = (int)b->ctx.size();
  // fnLoop walks contiguous virtual-block ranges (blocks[k] workgroups);
  // fnOne runs one workgroup per virtual block (blocks[k] == 0)
  using Fn = void (*)(BatchLaunchT<R>);
  auto launch = [&](int k, Fn fnLoop, Fn fnOne, int variant, unsigned block = kBlock) {
    BatchLaunchT<R> bl{static_cast<const KPT<R>*>(b->dTable), b->dStart + (size_t)k * (n + 1),
                       b->dGx + (size_t)k * n, n, variant, b->total[k], b->dOwner + b->ownerOff[k]};
    const bool one = b->blocks[k] == 0 || b->blocks[k] >= b->total[k];
    const unsigned grid = one ? b->total[k] : b->blocks[k];
    const Fn fn = one ? fnOne : fnLoop;
    if (timing) (void)hipEventRecord(b->ev[2 * k], b->stream);
    hipLaunchKernelGGL(fn, dim3(grid), dim3(block), 0, b->stream, bl);
    if (timing) (void)hipEventRecord(b->ev[2 * k + 1], b->stream);
  };
  launch(BK_DT, k_dt_reduce_batch<R, false>, k_dt_reduce_batch<R, true>, cur);
  {
    BatchLaunchT<R> bl{static_cast<const KPT<R>*>(b->dTable), nullptr, nullptr, n, cur, 0u, nullptr};
    hipLaunchKernelGGL(k_dt_finalize_batch<R>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, b->stream, bl);
  }
  launch(BK_NODE, k_lagrange_node_batch<R, false>, k_lagrange_node_batch<R, true>, cur);
  static const Fn cellLoop[3] = {k_lagrange_cell_batch<R, 1, false>, k_lagrange_cell_batch<R, 2, false>,
                                 k_lagrange_cell_batch<R, 3, false>};
  static const Fn cellOne[3] = {k_lagrange_cell_batch<R, 1, true>, k_lagrange_cell_batch<R, 2, true>,
                                k_lagrange_cell_batch<R, 3, true>};
  launch(BK_CELL, cellLoop[b->m - 1], cellOne[b->m - 1], cur);
  for (int sweep = 1; sweep <= 2; ++sweep) {
    const bool xd = (sweep == 1) ? xty : !xty;
    const int fuse = (sweep == 2) ? 1 : 0;
    const int v = cur + 2 * (sweep - 1) + 4 * (xd ? 1 : 0);
    if (b->o3)
      launch(xd ? BK_SWX : BK_SWY, o3BatchKernel<R, false>(xd, b->m), o3BatchKernel<R, true>(xd, b->m), v + 8 * fuse);
    else if (!xd && b->walkY)
      launch(BK_SWY, walkBatchKernel<R, false>(b->m, b->pp), walkBatchKernel<R, true>(b->m, b->pp), v + 8 * fuse);
    else
      launch(xd ? BK_SWX : BK_SWY, sweepBatchKernel<R, false>(xd, b->m, b->pp),
             sweepBatchKernel<R, true>(xd, b->m, b->pp), v + 8 * fuse, sweepWaveBlock(xd, b->mixed, b->pp));
    if (xd) launch(BK_DUX, k_dual_wave_batch<R, true, false>, k_dual_wave_batch<R, true, true>, v);
    else if (b->dualWaveY)
      launch(BK_DUY, k_dual_wave_batch<R, false, false>, k_dual_wave_batch<R, false, true>, v, dualWaveBlock(false, b->mixed));
    else launch(BK_DUY, k_dual_batch<R, false, false>, k_dual_batch<R, false, true>, v);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MHY_OK : batchFail(b, MHY_ERR_DEVICE, std::string("batch launch: ") + hipGetErrorString(e));
}

int batchEnqueueStep(mhy_batch* b, int cur, bool xty, bool timing) {
  return b->mixed ? batchEnqueueStepR<float>(b, cur, xty, timing) : batchEnqueueStepR<double>(b, cur, xty, timing);
}

// captures the two step parities of the lock-step ensemble as graphs:
// exec[c] is the step that starts with parity cur = c
int batchCapture(mhy_batch* b) {
  for (int c = 0; c < 2; ++c) {
    const bool xty = (c == b->cur0) ? b->xty0 : !b->xty0;
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return batchFail(b, MHY_ERR_DEVICE, "graph capture begin");
    int st = batchEnqueueStep(b, c, xty, false);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (st != MHY_OK) return st;
    if (e != hipSuccess) return batchFail(b, MHY_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&b->exec[c], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return batchFail(b, MHY_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
  }
  return MHY_OK;
}

// picks, per kernel, the candidate workgroup count with the shortest time
void batchPickBlocks(mhy_batch* b) {
  for (int k = 0; k < BK_COUNT; ++k) {
    int best = 0;
    for (int c = 1; c < kNCand; ++c)
      if (b->tuneMs[k][c] > 0.f && (b->tuneMs[k][best] <= 0.f || b->tuneMs[k][c] < b->tuneMs[k][best])) best = c;
    b->blocks[k] = kCand[best];
  }
}

}  // namespace

extern "C" {

int mhy_batch_create(const mhy_config* cfgs, int n, mhy_batch** out) {
  if (!cfgs || n <= 0 || !out) return MHY_ERR_ARG;
  *out = nullptr;
  mhy_batch* b = new mhy_batch();
  *out = b;
  for (int i = 0; i < n; ++i) {
    const mhy_config& c = cfgs[i];
    if (c.schema != MHY_SCHEMA_VNR || !c.avec_projection || c.sans_lagrange)
      return batchFail(b, MHY_ERR_ARG, "batch problems must use the VNR Lagrange + remap mode");
    if (i > 0 && (c.nbmat != cfgs[0].nbmat || c.pp != cfgs[0].pp || c.precision != cfgs[0].precision ||
                  (c.projection_order == 3) != (cfgs[0].projection_order == 3)))
      return batchFail(b, MHY_ERR_ARG,
                       "batch problems must share the material count, the PenteBorne option, the precision and "
                       "whether the projection is of order 3");
  }
  b->m = cfgs[0].nbmat;
  b->pp = cfgs[0].pp == 1;
  b->mixed = cfgs[0].precision == 1;
  b->o3 = cfgs[0].projection_order == 3;
  if (const char* dy = std::getenv("MAHYCO_DUAL_Y")) b->dualWaveY = (std::string(dy) != "node");
  if (const char* wy = std::getenv("MAHYCO_SWEEP_Y")) b->walkY = (std::string(wy) == "walk");
  if (const char* g = std::getenv("MAHYCO_BATCH_GRAPH")) b->useGraph = (std::string(g) != "0");
  if (const char* t = std::getenv("MAHYCO_BATCH_TUNE")) b->tune = (std::string(t) != "0");
  b->ctx.resize(n, nullptr);
  for (int i = 0; i < n; ++i) {
    int st = createContext(&cfgs[i], &b->ctx[i], true);
    if (st != MHY_OK)
      return batchFail(b, st, "problem " + std::to_string(i) + ": " + (b->ctx[i] ? mhy_last_error(b->ctx[i]) : ""));
    b->ctx[i]->sweepWave = true;
    b->ctx[i]->fuseFinal = true;
    b->cells += b->ctx[i]->ncell;
  }
  b->cur0 = b->cur = b->ctx[0]->cur;
  b->xty0 = b->xty = b->ctx[0]->x_then_y_n;
  if (hipStreamCreate(&b->stream) != hipSuccess) return batchFail(b, MHY_ERR_DEVICE, "stream");
  for (auto& e : b->ev)
    if (hipEventCreate(&e) != hipSuccess) return batchFail(b, MHY_ERR_DEVICE, "events");
  // launch parameters of every problem and step variant (KPT<float> has the
  // same layout, so the mixed form uploads the same entries)
  static_assert(sizeof(KPT<float>) == sizeof(KP), "parameter blocks differ");
  std::vector<KP> table((size_t)kVariants * n);
  for (int v = 0; v < kVariants; ++v)
    for (int i = 0; i < n; ++i) {
      KP kp = makeKP(b->ctx[i]);
      kp.cur = v & 1;
      kp.nxt = 1 - kp.cur;
      kp.sweep = ((v >> 1) & 1) + 1;
      kp.xdir = (v >> 2) & 1;
      kp.fuseFinal = (v >> 3) & 1;
      table[(size_t)v * n + i] = kp;
    }
  // virtual-block layout of every kernel
  std::vector<unsigned> start((size_t)BK_COUNT * (n + 1)), gx((size_t)BK_COUNT * n);
  for (int k = 0; k < BK_COUNT; ++k) {
    unsigned acc = 0;
    for (int i = 0; i < n; ++i) {
      const mhy_ctx* x = b->ctx[i];
      const long nx = x->o.nx, ny = x->o.ny;
      dim3 g;
      switch (k) {
        case BK_DT:  // a few cells per thread: fewer, longer virtual blocks
          g = dim3((unsigned)std::min<long>(kReduceBlocks, (x->ncell + 4 * kBlock - 1) / (4 * kBlock)));
          break;
        case BK_NODE: g = gridFor(x->nnode); break;
        case BK_CELL: g = gridFor(x->ncell); break;
        case BK_SWX: g = b->o3 ? sweepO3Grid(true, nx, ny) : sweepWaveGrid(true, nx, ny, b->mixed, b->pp); break;
        case BK_SWY:
          g = b->o3      ? sweepO3Grid(false, nx, ny)
              : b->walkY ? sweepWalkGrid(nx, ny)
                         : sweepWaveGrid(false, nx, ny, b->mixed, b->pp);
          break;
        case BK_DUX: g = dualWaveGrid(true, nx, ny, b->mixed); break;
        case BK_DUY: g = b->dualWaveY ? dualWaveGrid(false, nx, ny, b->mixed) : gridFor(x->nnode); break;
        default: g = gridFor(x->nnode); break;  // per-node dual along y
      }
      start[(size_t)k * (n + 1) + i] = acc;
      gx[(size_t)k * n + i] = g.x;
      acc += g.x * g.y;
    }
    start[(size_t)k * (n + 1) + n] = acc;
    b->total[k] = acc;
  }
  std::vector<unsigned> owner;
  for (int k = 0; k < BK_COUNT; ++k) {
    b->ownerOff[k] = owner.size();
    for (int i = 0; i < n; ++i)
      owner.insert(owner.end(), start[(size_t)k * (n + 1) + i + 1] - start[(size_t)k * (n + 1) + i], (unsigned)i);
  }
  if (hipMalloc(&b->dTable, table.size() * sizeof(KP)) != hipSuccess ||
      hipMalloc(&b->dStart, start.size() * sizeof(unsigned)) != hipSuccess ||
      hipMalloc(&b->dGx, gx.size() * sizeof(unsigned)) != hipSuccess ||
      hipMalloc(&b->dOwner, std::max<size_t>(owner.size(), 1) * sizeof(unsigned)) != hipSuccess)
    return batchFail(b, MHY_ERR_DEVICE, "batch tables");
  if (hipMemcpy(b->dTable, table.data(), table.size() * sizeof(KP), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->dStart, start.data(), start.size() * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->dGx, gx.data(), gx.size() * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->dOwner, owner.data(), owner.size() * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess)
    return batchFail(b, MHY_ERR_DEVICE, "batch tables upload");
  if (!b->tune) b->tuneStep = kTuneSteps;  // keep one workgroup per virtual block
  return MHY_OK;
}

int mhy_batch_destroy(mhy_batch* b) {
  if (!b) return MHY_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto& e : b->exec)
    if (e) (void)hipGraphExecDestroy(e);
  for (auto& e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->dTable) (void)hipFree(b->dTable);
  if (b->dStart) (void)hipFree(b->dStart);
  if (b->dGx) (void)hipFree(b->dGx);
  if (b->dOwner) (void)hipFree(b->dOwner);
  for (mhy_ctx* x : b->ctx) mhy_destroy(x);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return MHY_OK;
}

const char* mhy_batch_last_error(const mhy_batch* b) { return b ? b->err.c_str() : "null batch"; }

long long mhy_batch_cells(const mhy_batch* b) { return b ? b->cells : 0; }

int mhy_batch_context(mhy_batch* b, int i, mhy_ctx** ctx) {
  if (!b || !ctx || i < 0 || i >= (int)b->ctx.size()) return MHY_ERR_ARG;
  *ctx = b->ctx[i];
  return MHY_OK;
}

int mhy_batch_run(mhy_batch* b, int nsteps) {
  if (!b || nsteps < 0) return MHY_ERR_ARG;
  if (b->ctx.empty() || !b->dTable) return MHY_ERR_ARG;
  for (int s = 0; s < nsteps; ++s) {
    int st;
    if (b->tuneStep < kTuneSteps) {
      // two consecutive steps (both sweep orders: a sweep kernel is sweep 1
      // in one and the heavier fused sweep 2 in the other) with every kernel
      // at candidate tuneStep / 2, timed
      const int cand = b->tuneStep / 2;
      for (int k = 0; k < BK_COUNT; ++k) b->blocks[k] = kCand[cand];
      if ((st = batchEnqueueStep(b, b->cur, b->xty, true)) != MHY_OK) return st;
      if (hipStreamSynchronize(b->stream) != hipSuccess) return batchFail(b, MHY_ERR_DEVICE, "batch sync");
      const bool verbose = std::getenv("MAHYCO_BATCH_VERBOSE") != nullptr;
      for (int k = 0; k < BK_COUNT; ++k) {  // every kernel runs once per step
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, b->ev[2 * k], b->ev[2 * k + 1]) == hipSuccess) b->tuneMs[k][cand] += ms;
        if (verbose) std::fprintf(stderr, "[batch tune] kernel %d blocks %u: %.4f ms\n", k, kCand[cand], ms);
      }
      if (++b->tuneStep == kTuneSteps) {
        batchPickBlocks(b);
        if (verbose)
          for (int k = 0; k < BK_COUNT; ++k) std::fprintf(stderr, "[batch tune] kernel %d -> %u\n", k, b->blocks[k]);
      }
    } else if (b->useGraph) {
      if (!b->exec[0] && (st = batchCapture(b)) != MHY_OK) return st;
      if (hipGraphLaunch(b->exec[b->cur], b->stream) != hipSuccess)
        return batchFail(b, MHY_ERR_DEVICE, "graph launch");
    } else {
      if ((st = batchEnqueueStep(b, b->cur, b->xty, false)) != MHY_OK) return st;
    }
    b->cur = 1 - b->cur;
    b->xty = !b->xty;
    ++b->iteration;
  }
  if (hipStreamSynchronize(b->stream) != hipSuccess) return batchFail(b, MHY_ERR_DEVICE, "batch sync");
  // hand the state back to the problem contexts and collect their statuses
  int status = MHY_OK;
  for (size_t i = 0; i < b->ctx.size(); ++i) {
    mhy_ctx* x = b->ctx[i];
    x->cur = b->cur;
    x->x_then_y_n = b->xty;
    x->iteration = (int)b->iteration;
    Scalars sc;
    if (hipMemcpy(&sc, x->f.sc, sizeof(Scalars), hipMemcpyDeviceToHost) != hipSuccess)
      return batchFail(b, MHY_ERR_DEVICE, "batch scalars");
    if (sc.err && status == MHY_OK) {
      status = sc.err;
      b->err = "problem " + std::to_string(i) + ": kernel status " + std::to_string(sc.err);
      x->status = sc.err;
    }
  }
  return status;
}

}  // extern "C"
