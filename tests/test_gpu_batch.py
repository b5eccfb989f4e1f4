"""Ensembles on the GPU (mhy_batch_*): every problem of a batch must end
bitwise equal to the same problem stepped alone, whether the steps run from
the tuning launches, from the captured HIP graphs (replayed several times
with other GPU work and synchronisations in between) or eagerly."""
import os

import numpy as np
import pytest

import mahyco_amd
from mahyco_amd import configs

from gpu_helpers import log

pytestmark = pytest.mark.gpu
FIELDS = ["density", "pressure", "energy", "sound_speed", "velocity_x", "velocity_y", "node_mass"]


def _compare(batch, singles, cfgs, what):
    for i, c in enumerate(cfgs):
        pb, ref = batch.problem(i), singles[i]
        for f in FIELDS:
            a, r = pb.get(f), ref.get(f)
            assert np.array_equal(a, r, equal_nan=True), f"{what}: problem {i} ({c.nx}x{c.ny}) {f}"
        for k in range(c.nbmat):
            assert np.array_equal(pb.get("fracvol", k), ref.get("fracvol", k)), f"{what}: problem {i} fracvol{k}"
        ta, tr = pb.time_state(), ref.time_state()
        assert (ta.t_n, ta.deltat_n) == (tr.t_n, tr.deltat_n), f"{what}: problem {i} time"


@pytest.mark.parametrize("graphs", [True, False], ids=["graphs", "eager"])
def test_batch_matches_single_problems(graphs):
    cfgs = configs.ensemble(n=9, side_min=12, side_max=96)
    old = os.environ.get("MAHYCO_BATCH_GRAPH")
    os.environ["MAHYCO_BATCH_GRAPH"] = "1" if graphs else "0"
    try:
        batch = mahyco_amd.Batch(cfgs)
    finally:
        if old is None:
            del os.environ["MAHYCO_BATCH_GRAPH"]
        else:
            os.environ["MAHYCO_BATCH_GRAPH"] = old
    singles = [mahyco_amd.Context(c) for c in cfgs]
    try:
        log(f"[batch] {len(cfgs)} problems, {batch.cells} cells, graphs={graphs}")
        done = 0
        for chunk in (3, 4, 5, 3):  # 10 tuning steps, then replays
            batch.run(chunk)
            for s in singles:
                s.step(chunk)
            done += chunk
            _compare(batch, singles, cfgs, f"after {done} steps")
    finally:
        for s in singles:
            s.close()
        batch.close()


def test_batch_of_one_equals_the_problem():
    c = configs.riemann(40, 0.45, 1.0, 1.0, 0.125, 0.1, 1.4, 1.67)
    batch = mahyco_amd.Batch([c])
    single = mahyco_amd.Context(c)
    try:
        batch.run(8)
        single.step(8)
        _compare(batch, [single], [c], "one problem")
    finally:
        single.close()
        batch.close()


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_third_order_batch_matches_single_problems(precision):
    """Ensembles at projection order 3 (LDS-tile sweeps) against the
    problems alone.  About 6000 tiles per sweep, more than the 1024 to 8192
    workgroups the tuning steps try, so those steps run the range-walking
    kernels (several tiles per workgroup, the LDS tile reused) as well as
    one workgroup per tile."""
    cfgs = configs.ensemble(n=7, side_min=300, side_max=700, seed=77)
    for c in cfgs:
        c.projection_order = 3
        c.precision = precision
    batch = mahyco_amd.Batch(cfgs)
    singles = [mahyco_amd.Context(c) for c in cfgs]
    try:
        log(f"[batch] order 3, {len(cfgs)} problems, {batch.cells} cells, precision={precision}")
        done = 0
        for chunk in (4, 6, 3):  # the tuning steps, then graph replays
            batch.run(chunk)
            for s in singles:
                s.step(chunk)
            done += chunk
            _compare(batch, singles, cfgs, f"after {done} steps")
    finally:
        for s in singles:
            s.close()
        batch.close()
