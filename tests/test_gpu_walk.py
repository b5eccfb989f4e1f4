"""The two kernels of the cell sweeps along y must give bitwise the same
state: the column-walking kernel (one lane per column, rows streamed through
registers) and the wave-strip kernel (4 columns x 16 rows per wave with
two-cell halos).  MAHYCO_SWEEP_Y forces either; the choice is read when a
context or a batch is created.  Covers the FP64 and the mixed forms, every
material count, the PenteBorne variants and meshes whose sides are not
multiples of the walker's 64 columns and 32 rows, plus the batched walker."""
import os
from contextlib import contextmanager

import numpy as np
import pytest

import mahyco_amd
from mahyco_amd import configs

from gpu_helpers import field_names, log
from test_gpu_parity import _small, extension_cases

pytestmark = pytest.mark.gpu
STEPS = 12


@contextmanager
def sweep_y(mode):
    old = os.environ.get("MAHYCO_SWEEP_Y")
    os.environ["MAHYCO_SWEEP_Y"] = mode
    try:
        yield
    finally:
        if old is None:
            del os.environ["MAHYCO_SWEEP_Y"]
        else:
            os.environ["MAHYCO_SWEEP_Y"] = old


def _cases():
    out = {
        "C4_100x70": configs.triple_layout(100, 70),
        "riemann_130": configs.riemann(130, 0.45, 1.0, 1.0, 0.125, 0.1, 1.4, 1.67),
        "C3_200": _small(configs.point_explosion(), 200),
        "C2_pp_150": _small(configs.riemann_perturbed(), 150),
    }
    # the third-order remap has its own per-cell sweep (no walker)
    out.update({k: c for k, c in extension_cases().items() if c.projection_order != 3})
    return out


CASES = _cases()


def _differing(a, b, cfg):
    return [f"{n}[{k}]" for n, k in field_names(cfg) if not np.array_equal(a.get(n, k), b.get(n, k), equal_nan=True)]


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name", list(CASES))
def test_walker_equals_wave_strip(name, precision):
    cfg = CASES[name]
    cfg.precision = precision
    try:
        log(f"[walk] {name} {cfg.nx}x{cfg.ny} m={cfg.nbmat} precision={precision}")
        with sweep_y("wave"):
            wave = mahyco_amd.Context(cfg)
        with sweep_y("walk"):
            walk = mahyco_amd.Context(cfg)
        try:
            for chunk in (1, 1, STEPS - 2):
                wave.step(chunk)
                walk.step(chunk)
                diff = _differing(walk, wave, cfg)
                assert not diff, f"{name}: walker differs in {diff}"
                ta, tb = walk.time_state(), wave.time_state()
                assert (ta.t_n, ta.deltat_n) == (tb.t_n, tb.deltat_n)
        finally:
            wave.close()
            walk.close()
    finally:
        cfg.precision = 0


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_batched_walker_equals_single_wave_strip(precision):
    cfgs = configs.ensemble(n=7, side_min=12, side_max=140)
    for c in cfgs:
        c.precision = precision
    with sweep_y("walk"):
        batch = mahyco_amd.Batch(cfgs)
    with sweep_y("wave"):
        singles = [mahyco_amd.Context(c) for c in cfgs]
    try:
        log(f"[walk] batch of {len(cfgs)} problems, precision={precision}")
        for chunk in (4, 6, 3):  # the tuning steps, then graph replays
            batch.run(chunk)
            for s in singles:
                s.step(chunk)
            for i, c in enumerate(cfgs):
                diff = _differing(batch.problem(i), singles[i], c)
                assert not diff, f"problem {i} ({c.nx}x{c.ny}) differs in {diff}"
    finally:
        for s in singles:
            s.close()
        batch.close()


@contextmanager
def lag_y_separate():
    old = os.environ.get("MAHYCO_LAG_Y")
    os.environ["MAHYCO_LAG_Y"] = "separate"
    try:
        yield
    finally:
        if old is None:
            del os.environ["MAHYCO_LAG_Y"]
        else:
            os.environ["MAHYCO_LAG_Y"] = old


STORE_CASES = {k: c for k, c in CASES.items() if c.nbmat == 1 and c.pp != 1}
STORE_CASES["C3_37x301"] = _small(configs.point_explosion(), 37, 301)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name", list(STORE_CASES))
def test_first_sweep_walker_equals_wave_strip(name, precision):
    """With the Lagrange cell update as its own kernel on the y-first steps
    (the tuner's choice on the 4096^2 point explosion), their first sweep
    along y runs on the walker too, reading the stored remap vector: bitwise
    the wave strip's state."""
    cfg = STORE_CASES[name]
    cfg.precision = precision
    try:
        log(f"[walk-first] {name} {cfg.nx}x{cfg.ny} precision={precision}")
        with lag_y_separate():
            with sweep_y("wave"):
                wave = mahyco_amd.Context(cfg)
            with sweep_y("walk"):
                walk = mahyco_amd.Context(cfg)
        try:
            for chunk in (1, 1, STEPS - 2):
                wave.step(chunk)
                walk.step(chunk)
                diff = _differing(walk, wave, cfg)
                assert not diff, f"{name}: walker differs in {diff}"
                ta, tb = walk.time_state(), wave.time_state()
                assert (ta.t_n, ta.deltat_n) == (tb.t_n, tb.deltat_n)
        finally:
            wave.close()
            walk.close()
    finally:
        cfg.precision = 0
