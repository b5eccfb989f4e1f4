// Kernel entry points of the hot path (defined in lagrange.hip,
// sweep_wave.hip, sweep_cell.hip, remap.hip).  R is the storage type of the
// bulk fields: double (FP64 form) or float (mixed form).
#pragma once

#include "common.hpp"

namespace mhy {

// Eucclhyd Lagrange phase and final remap (eucclhyd.hip); M: materials
template <int M>
__global__ void k_euc_start(KP kp, EucFields e);
__global__ void k_euc_dt_finalize(KP kp);
template <int M>
__global__ void k_euc_eos(KP kp);
template <int M>
__global__ void k_euc_node(KP kp, EucFields e);
template <int M>
__global__ void k_euc_cell(KP kp, EucFields e);
template <int M>
__global__ void k_euc_final(KP kp, EucFields e);

template <class R>
__global__ void k_dt_reduce(KPT<R> kp);
template <class R>
__global__ void k_dt_finalize(KPT<R> kp);
template <class R>
__global__ void k_dt_finalize_batch(BatchLaunchT<R> bl);
template <class R>
__global__ void k_dt_hold(KPT<R> kp);
template <class R>
__global__ void k_lagrange_node(KPT<R> kp);
// the same node update on the fixed Euler grid, interior and boundary nodes
template <class R>
__global__ void k_lagrange_node_inner(KPT<R> kp);
template <class R>
__global__ void k_lagrange_node_edge(KPT<R> kp);
template <class R, int M>
__global__ void k_lagrange_cell(KPT<R> kp);
template <class R, bool ONE>
__global__ void k_dt_reduce_batch(BatchLaunchT<R> bl);
template <class R, bool ONE>
__global__ void k_lagrange_node_batch(BatchLaunchT<R> bl);
template <class R, int M, bool ONE>
__global__ void k_lagrange_cell_batch(BatchLaunchT<R> bl);
template <class R, int M, bool PP, bool XD, bool ONE>
__global__ void k_sweep_wave_batch(BatchLaunchT<R> bl);
template <class R, bool XD, bool ONE>
__global__ void k_dual_wave_batch(BatchLaunchT<R> bl);
template <class R, bool XD, bool ONE>
__global__ void k_dual_batch(BatchLaunchT<R> bl);
// per-cell stencil sweeps: FP64 form only
template <int M, bool PP, bool XD>
__global__ void k_sweep_cell(KP kp);
// third-order sweep on LDS tiles (256-thread blocks)
template <class R, int M, bool XD>
__global__ void k_sweep_o3(KPT<R> kp);
template <class R, int M, bool XD, bool ONE>
__global__ void k_sweep_o3_batch(BatchLaunchT<R> bl);
dim3 sweepO3Grid(bool xd, long nx, long ny);
template <class R, int M, bool PP, bool XD>
__global__ void k_sweep_wave(KPT<R> kp);
// the first sweep with the cell part of the Lagrange phase fused in
template <class R, int M, bool PP, bool XD>
__global__ void k_sweep_wave_lag(KPT<R> kp);
dim3 sweepWaveGrid(bool xd, long nx, long ny, bool mixed, bool pp);
unsigned sweepWaveBlock(bool xd, bool mixed, bool pp);
// y sweeps with one lane per column walking down a band of rows
template <class R, int M, bool PP>
__global__ void k_sweep_walk(KPT<R> kp);
template <class R, int M, bool PP, bool ONE>
__global__ void k_sweep_walk_batch(BatchLaunchT<R> bl);
dim3 sweepWalkGrid(long nx, long ny);
template <class R, bool XD>
__global__ void k_dual(KPT<R> kp);
template <class R, bool XD>
__global__ void k_dual_wave(KPT<R> kp);
dim3 dualWaveGrid(bool xd, long nx, long ny, bool mixed);
// nodal remap along y with one lane per node column walking down a band
template <class R>
__global__ void k_dual_walk(KPT<R> kp);
// nodal remap along y on LDS tiles of 16 node columns x ROWS rows (32, 48, 64)
template <class R, int ROWS>
__global__ void k_dual_tile(KPT<R> kp);
dim3 dualTileGrid(long nx, long ny, int rows);
unsigned dualTileBlock();
dim3 dualWalkGrid(long nx, long ny);
unsigned dualWaveBlock(bool xd, bool mixed);
template <class R, int M>
__global__ void k_remap_final(KPT<R> kp);
template <class R>
__global__ void k_node_mass(KPT<R> kp);
template <class R>
__global__ void k_conservation(KPT<R> kp, double* out);
template <class R>
__global__ void k_env_expand(KPT<R> kp);

}  // namespace mhy
