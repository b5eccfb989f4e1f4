"""Shared helpers of the GPU parity tests: run the HIP path and the CPU
oracle side by side and compare every field of the state."""
import os
import sys

import numpy as np

import mahyco_amd
from mahyco_amd import abi
from oracle import refcpu

CELL = ["density", "pressure", "energy", "pseudo", "sound_speed", "divu", "cell_mass"]
MAT = ["fracvol", "mass_fraction", "density_env", "energy_env", "pressure_env", "pseudo_env", "mass_env",
       "sound_speed_env"]
NODE = ["velocity_x", "velocity_y", "node_mass", "coord_x", "coord_y", "out_vx", "out_vy"]


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def field_names(cfg):
    names = [(n, 0) for n in CELL] + [(n, 0) for n in NODE]
    names += [(n, k) for k in range(cfg.nbmat) for n in MAT]
    return names


def rel_linf(a, b):
    """max |a-b| / max(|b|, tiny), NaN-aware (NaN at the same place is equal)."""
    both_nan = np.isnan(a) & np.isnan(b)
    d = np.where(both_nan, 0.0, np.abs(a - b))
    if np.any(np.isnan(d)):
        return np.inf
    scale = max(np.max(np.abs(np.where(np.isnan(b), 0.0, b))), 1e-300)
    return float(np.max(d) / scale) if d.size else 0.0


def compare_state(gpu, orc, cfg, tol=1e-10):
    """Returns (worst relative Linf, name of that field, all bitwise equal)."""
    worst, worst_name, bitwise = 0.0, "", True
    for name, k in field_names(cfg):
        a, b = gpu.get(name, k), orc.get(name, k)
        same = np.array_equal(a, b, equal_nan=True)
        bitwise &= same
        if not same:
            r = rel_linf(a, b)
            if r > worst:
                worst, worst_name = r, f"{name}[{k}]"
    return worst, worst_name, bitwise


def run_pair(cfg, steps, checkpoints=(1,), tol=1e-10):
    """Steps GPU and oracle together, comparing at checkpoints and the end."""
    gpu = mahyco_amd.Context(cfg)
    orc = refcpu.Oracle(cfg)
    results = []
    try:
        w, n, bit = compare_state(gpu, orc, cfg)
        results.append((0, w, n, bit))
        assert w <= tol, f"initial state differs: {n} rel {w}"
        done = 0
        for target in sorted(set([c for c in checkpoints if c <= steps] + [steps])):
            gpu.step(target - done)
            st = orc.step(target - done)
            assert st == abi.MHY_OK, f"oracle status {st}"
            done = target
            w, n, bit = compare_state(gpu, orc, cfg)
            tg, to = gpu.time_state(), orc.time_state()
            results.append((target, w, n, bit))
            assert to.t_n == tg.t_n and to.deltat_n == tg.deltat_n, (to.t_n, tg.t_n, to.deltat_n, tg.deltat_n)
            assert w <= tol, f"step {target}: {n} rel Linf {w:.3e}"
    finally:
        gpu.close()
        orc.close()
    return results
