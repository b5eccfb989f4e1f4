"""The kernel variants that the environment selects must give bitwise the
state of the default kernels: the per-cell stencil sweeps
(MAHYCO_SWEEP_KERNEL=cell, FP64 form only), the separate final-remap kernel
(MAHYCO_FUSE_FINAL=0), the separate Lagrange cell kernel instead of its
fusion into the first sweep (MAHYCO_FUSE_LAGRANGE=0), the per-node dual
remap along y (MAHYCO_DUAL_Y=node,
FP64 form: in the mixed form it keeps FP64 velocity arithmetic) and the
column-walking and LDS-tile dual remaps along y (MAHYCO_DUAL_Y=walk,
tile, and the 32- and 64-row tiles tile32, tile64), the Lagrange volume and centroid formed again in the second sweep or
stored by the first (MAHYCO_SWEEP_GEOM=recompute, load), one node
kernel for interior and boundary nodes (MAHYCO_SPLIT_NODE=0), and a
one-material context storing its material fields in every cell
(MAHYCO_COMPACT=0), the separate Lagrange cell kernel on the steps whose first
sweep runs along y only (MAHYCO_LAG_Y=separate), in both storage forms.  The choice is read when a
context is created."""
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
            "separate_lagrange": ("MAHYCO_FUSE_LAGRANGE", "0"),
            "node_dual_y": ("MAHYCO_DUAL_Y", "node"), "walk_dual_y": ("MAHYCO_DUAL_Y", "walk"),
            "wave_dual_y": ("MAHYCO_DUAL_Y", "wave"), "tile_dual_y": ("MAHYCO_DUAL_Y", "tile"),
            "tile32_dual_y": ("MAHYCO_DUAL_Y", "tile32"), "tile64_dual_y": ("MAHYCO_DUAL_Y", "tile64"),
            "recomputed_geometry": ("MAHYCO_SWEEP_GEOM", "recompute"),
            "stored_geometry": ("MAHYCO_SWEEP_GEOM", "load"), "one_node_kernel": ("MAHYCO_SPLIT_NODE", "0"),
            "stored_material_fields": ("MAHYCO_COMPACT", "0"),
            "separate_lagrange_y": ("MAHYCO_LAG_Y", "separate")}
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


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("variant", sorted(VARIANTS))
@pytest.mark.parametrize("name", NAMES)
def test_variant_equals_default(name, variant, precision):
    if precision == 1 and variant == "cell_sweeps":
        pytest.skip("the per-cell stencil sweeps are FP64 only (the mixed form ignores the switch)")
    if precision == 1 and variant == "node_dual_y":
        pytest.skip("the per-node dual kernel keeps FP64 velocity arithmetic in the mixed form, "
                    "the wave and walking kernels FP32: equal only to tolerance")
    cfg = extension_cases()[name]
    cfg.precision = precision
    var, value = VARIANTS[variant]
    log(f"[variant] {variant} {name} {cfg.nx}x{cfg.ny} m={cfg.nbmat} precision={precision}")
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
