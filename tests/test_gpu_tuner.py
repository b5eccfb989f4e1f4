"""The launch-shape tuner: on a mesh of at least 2^20 cells the first steps
time the interchangeable y-direction kernel variants (nodal remap on LDS
tiles of 48, 32 or 64 node rows, wave strips or walking bands; cell sweep on wave strips or walking
bands; the Lagrange update of y-first steps fused or separate) and keep the
fastest.  Results must be bitwise those of the same problem with the tuner
off, through the tuning steps, the switch to the kept variants and the
graph replays that follow, in both storage forms."""
import os
from contextlib import contextmanager

import numpy as np
import pytest

import mahyco_amd
from mahyco_amd import configs

from gpu_helpers import field_names, log

pytestmark = pytest.mark.gpu


@contextmanager
def env(name, value):
    old = os.environ.get(name)
    os.environ[name] = value
    try:
        yield
    finally:
        if old is None:
            del os.environ[name]
        else:
            os.environ[name] = old


def _cases():
    a = configs.riemann(1024, 0.45, 1.0, 1.0, 0.125, 0.1, 1.4, 1.67)  # limited slopes: three sites
    b = configs.riemann_perturbed(n=1024)                              # PenteBorne: no fusion site
    return {"riemann_1024_minmod": a, "riemann_1024_pentborne": b}


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name", sorted(_cases()))
def test_tuned_run_equals_untuned(name, precision):
    cfg = _cases()[name]
    cfg.precision = precision
    log(f"[tuner] {name} precision={precision}")
    tuned = mahyco_amd.Context(cfg)
    with env("MAHYCO_AUTOTUNE", "0"):
        plain = mahyco_amd.Context(cfg)
    try:
        assert tuned.tuned_kernels() == ""  # still tuning
        done = 0
        for chunk in (3, 33, 6):  # tuning steps (at most 36), then graph replays
            tuned.run(chunk)
            plain.run(chunk)
            done += chunk
            diff = [f"{n}[{k}]" for n, k in field_names(cfg)
                    if not np.array_equal(tuned.get(n, k), plain.get(n, k), equal_nan=True)]
            assert not diff, f"after {done} steps: {diff}"
        choice = tuned.tuned_kernels()
        log(f"[tuner] {name} precision={precision}: {choice}")
        assert "dual_y=" in choice and "sweep_y=" in choice
        assert ("lagrange_y=" in choice) == (cfg.pp != 1)
    finally:
        tuned.close()
        plain.close()


def test_small_meshes_are_not_tuned():
    cfg = configs.riemann(64, 0.45, 1.0, 1.0, 0.125, 0.1, 1.4, 1.67)
    ctx = mahyco_amd.Context(cfg)
    try:
        ctx.run(20)
        assert ctx.tuned_kernels() == ""
    finally:
        ctx.close()
