"""The kernel variants that the environment selects must give bitwise the
state of the default kernels (FP64 form): the per-cell stencil sweeps
(MAHYCO_SWEEP_KERNEL=cell), the separate final-remap kernel
(MAHYCO_FUSE_FINAL=0) and the per-node dual remap along y
(MAHYCO_DUAL_Y=node).  The choice is read when a context is created."""
import os
from contextlib import contextmanager

import numpy as np
import pytest

import mahyco_amd

from gpu_helpers import field_names, log
from test_gpu_parity import extension_cases

pytestmark = pytest.mark.gpu
STEPS = 12
VARIANTS = {"cell_sweeps": ("MAHYCO_SWEEP_KERNEL", "cell"), "separate_final": ("MAHYCO_FUSE_FINAL", "0"),
            "node_dual_y": ("MAHYCO_DUAL_Y", "node")}
NAMES = ["C2_pp_superbee", "C2_pp_PB", "C2_pp_mixte_arithmeticG", "C3_point_explosion", "C4_triple_layout",
         "riemann_order1", "riemann_superBeeG_main", "unit_test_boundary_flux", "riemann_order3"]


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


@pytest.mark.parametrize("variant", sorted(VARIANTS))
@pytest.mark.parametrize("name", NAMES)
def test_variant_equals_default(name, variant):
    cfg = extension_cases()[name]
    var, value = VARIANTS[variant]
    log(f"[variant] {variant} {name} {cfg.nx}x{cfg.ny} m={cfg.nbmat}")
    ref = mahyco_amd.Context(cfg)
    with env(var, value):
        alt = mahyco_amd.Context(cfg)
    try:
        for chunk in (1, 1, STEPS - 2):
            ref.step(chunk)
            alt.step(chunk)
            diff = [f"{n}[{k}]" for n, k in field_names(cfg)
                    if not np.array_equal(alt.get(n, k), ref.get(n, k), equal_nan=True)]
            assert not diff, f"{variant} differs in {diff}"
    finally:
        ref.close()
        alt.close()
