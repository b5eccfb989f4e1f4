"""The 512-problem ensemble at its real size: two-material Riemann problems
on square meshes with sides drawn log-uniformly in 64..1024 (seed
20261204), stepped together as one batch (one launch per kernel over all
problems, captured HIP graphs after the tuning steps).

* FP64 form: a seeded sample of problems (the smallest, the largest and
  four between them by cell count) against the CPU oracle after the ten
  tuning steps and two graph replays, every field within 1e-10 relative
  Linf (in practice bitwise).
* Mixed FP32 / FP64 form: the same sample against the oracle after 200
  steps, relative L1 of density and energy <= 1e-5 (the precision-form
  target).
* Both forms over 200 steps, every problem: each material's mass conserved
  (relative drift <= 1e-10 mixed, 1e-13 FP64), volume fractions in [0, 1]
  summing to 1, density positive and every field finite.
"""
import numpy as np
import pytest

import mahyco_amd
from mahyco_amd import abi, configs, precision
from oracle import refcpu

from gpu_helpers import compare_state, log

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

STEPS = 200
L1_TOL = 1e-5
MASS_TOL = {0: 1e-13, 1: 1e-10}


def _ensemble(prec):
    cfgs = configs.ensemble()
    for c in cfgs:
        c.precision = prec
    return cfgs


def _sample(cfgs, n_between=4, seed=7):
    """Indices of the smallest and the largest problem and n_between others
    drawn from the rest."""
    cells = np.array([c.nx * c.ny for c in cfgs])
    lo, hi = int(np.argmin(cells)), int(np.argmax(cells))
    rest = [i for i in range(len(cfgs)) if i not in (lo, hi)]
    rng = np.random.default_rng(seed)
    mid = sorted(rng.choice(rest, n_between, replace=False).tolist(), key=lambda i: cells[i])
    return [lo] + mid + [hi]


def _invariants(batch, cfgs, mass0, prec, what):
    worst = 0.0
    for i, c in enumerate(cfgs):
        pb = batch.problem(i)
        sums = pb.conservation()
        for k in range(c.nbmat):
            m0 = mass0[i][2 + k]
            if m0 > 0:
                worst = max(worst, abs(sums[2 + k] - m0) / m0)
        fs = np.stack([pb.get("fracvol", k) for k in range(c.nbmat)])
        assert fs.min() >= 0.0 and fs.max() <= 1.0, f"{what}: problem {i} volume fraction outside [0, 1]"
        assert np.allclose(fs.sum(axis=0), 1.0, rtol=0, atol=1e-12), f"{what}: problem {i} fractions do not sum to 1"
        rho = pb.get("density")
        assert np.all(np.isfinite(rho)) and rho.min() > 0.0, f"{what}: problem {i} density"
        for fld in ("energy", "pressure", "velocity_x", "velocity_y"):
            assert np.all(np.isfinite(pb.get(fld))), f"{what}: problem {i} {fld} not finite"
    log(f"[ensemble] {what}: worst per-material mass drift {worst:.2e}")
    assert worst <= MASS_TOL[prec], f"{what}: mass drift {worst:.3e}"


def test_ensemble_fp64_sample_equals_oracle_then_invariants():
    cfgs = _ensemble(0)
    sample = _sample(cfgs)
    batch = mahyco_amd.Batch(cfgs)
    try:
        log(f"[ensemble] FP64: {len(cfgs)} problems, {batch.cells} cells; sample {sample}")
        mass0 = [batch.problem(i).conservation() for i in range(len(cfgs))]
        batch.run(12)  # 10 tuning steps, then graph replays
        for i in sample:
            c = cfgs[i]
            orc = refcpu.Oracle(c)
            try:
                assert orc.step(12) == abi.MHY_OK
                pb = batch.problem(i)
                tg, to = pb.time_state(), orc.time_state()
                assert (tg.t_n, tg.deltat_n) == (to.t_n, to.deltat_n), f"problem {i} time"
                worst, field, bitwise = compare_state(pb, orc, c)
                log(f"[ensemble] FP64 problem {i} ({c.nx}x{c.ny}): rel Linf {worst:.2e} ({field or '-'}), "
                    f"bitwise={bitwise}")
                assert worst <= 1e-10, f"problem {i} {field}: rel Linf {worst:.3e}"
            finally:
                orc.close()
        batch.run(STEPS - 12)
        _invariants(batch, cfgs, mass0, 0, f"FP64 after {STEPS} steps")
    finally:
        batch.close()


def test_ensemble_mixed_sample_against_oracle_then_invariants():
    cfgs = _ensemble(1)
    sample = _sample(cfgs)
    batch = mahyco_amd.Batch(cfgs)
    try:
        log(f"[ensemble] mixed: {len(cfgs)} problems, {batch.cells} cells; sample {sample}")
        mass0 = [batch.problem(i).conservation() for i in range(len(cfgs))]
        batch.run(STEPS)
        _invariants(batch, cfgs, mass0, 1, f"mixed after {STEPS} steps")
        for i in sample:
            c = cfgs[i]
            c.precision = 0
            orc = refcpu.Oracle(c)
            try:
                assert orc.step(STEPS) == abi.MHY_OK
                pb = batch.problem(i)
                ta, tb = orc.time_state(), pb.time_state()
                assert abs(ta.t_n - tb.t_n) <= 1e-5 * ta.t_n, f"problem {i} time {tb.t_n} vs {ta.t_n}"
                errs = precision.field_errors(pb, orc, c.nbmat)
                log(f"[ensemble] mixed problem {i} ({c.nx}x{c.ny}): " +
                    " ".join(f"{k} {v[0]:.1e}/{v[1]:.1e}" for k, v in errs.items()))
                for k in ("rho", "e"):
                    assert errs[k][0] <= L1_TOL, f"problem {i}: relative L1 of {k} = {errs[k][0]:.3e}"
            finally:
                orc.close()
    finally:
        batch.close()
