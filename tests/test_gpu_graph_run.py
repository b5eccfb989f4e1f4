"""mhy_run replays each step as a captured HIP graph (one per step parity
and sweep order).  A context advanced by mhy_run in chunks, with eager steps
and host reads in between, must stay bitwise equal to the same problem
stepped eagerly, replay after replay; with kernel timing on, mhy_run takes
the eager path and must agree as well."""
import pytest

import mahyco_amd

from gpu_helpers import field_names, log
from test_gpu_parity import extension_cases

pytestmark = pytest.mark.gpu
NAMES = ["C2_pp_superbee", "C3_point_explosion", "C4_triple_layout", "riemann_order3", "unit_test_boundary_flux",
         "C2_pp_PB"]


def _differing(a, b, cfg):
    import numpy as np
    return [f"{n}[{k}]" for n, k in field_names(cfg) if not np.array_equal(a.get(n, k), b.get(n, k), equal_nan=True)]


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name", NAMES)
def test_graph_run_equals_eager_steps(name, precision):
    cfg = extension_cases()[name]
    cfg.precision = precision
    log(f"[graph] {name} {cfg.nx}x{cfg.ny} precision={precision}")
    g = mahyco_amd.Context(cfg)
    e = mahyco_amd.Context(cfg)
    try:
        done = 0
        for kind, n in (("run", 5), ("step", 1), ("run", 3), ("run", 7), ("step", 2), ("run", 4)):
            if kind == "run":
                g.run(n)
            else:
                g.step(n)
            e.step(n)
            done += n
            diff = _differing(g, e, cfg)
            assert not diff, f"after {done} steps ({kind} {n}): {diff}"
            tg, te = g.time_state(), e.time_state()
            assert (tg.t_n, tg.deltat_n) == (te.t_n, te.deltat_n)
    finally:
        g.close()
        e.close()


def test_run_with_kernel_timing_is_eager_and_equal():
    cfg = extension_cases()["C2_pp_superbee"]
    g = mahyco_amd.Context(cfg)
    e = mahyco_amd.Context(cfg)
    try:
        g.run(4)  # graphs
        g.enable_timing(True)
        g.run(4)  # eager, with events
        stats = g.kernel_stats()
        g.enable_timing(False)
        g.run(3)  # graphs again
        e.step(11)
        # each direction is swept twice per step: the first sweep (with the
        # Lagrange cell update fused in, its own kernel id) and the second
        for d in ("x", "y"):
            n = sum(v["launches"] for k, v in stats.items() if k.endswith(f"sweep_{d}"))
            assert n == 4, stats
        diff = _differing(g, e, cfg)
        assert not diff, diff
    finally:
        g.close()
        e.close()
