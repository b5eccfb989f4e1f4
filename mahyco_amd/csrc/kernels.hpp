// Kernel entry points of the hot path (defined in lagrange.hip,
// sweep_wave.hip, sweep_cell.hip, remap.hip).
#pragma once

#include "common.hpp"

namespace mhy {

__global__ void k_dt_reduce(KP kp);
__global__ void k_dt_hold(KP kp);
__global__ void k_lagrange_node(KP kp);
__global__ void k_lagrange_cell(KP kp);
__global__ void k_dt_reduce_batch(BatchLaunch bl);
__global__ void k_lagrange_node_batch(BatchLaunch bl);
__global__ void k_lagrange_cell_batch(BatchLaunch bl);
template <int M, bool PP, bool XD>
__global__ void k_sweep_wave_batch(BatchLaunch bl);
template <bool XD>
__global__ void k_dual_wave_batch(BatchLaunch bl);
template <bool XD>
__global__ void k_dual_batch(BatchLaunch bl);
template <int M, bool PP, bool XD>
__global__ void k_sweep_cell(KP kp);
template <int M, bool PP, bool XD>
__global__ void k_sweep_wave(KP kp);
dim3 sweepWaveGrid(bool xd, long nx, long ny);
template <bool XD>
__global__ void k_dual(KP kp);
template <bool XD>
__global__ void k_dual_wave(KP kp);
dim3 dualWaveGrid(bool xd, long nx, long ny);
template <int M>
__global__ void k_remap_final(KP kp);
__global__ void k_node_mass(KP kp);
__global__ void k_conservation(KP kp, double* out);

}  // namespace mhy
