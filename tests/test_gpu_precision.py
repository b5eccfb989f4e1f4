"""Mixed FP32 / FP64 form on the GPU (precision 1): FP32 storage of the bulk
fields and FP32 slope / limiter / PenteBorne arithmetic, FP64 volume
fractions, masses, remap vectors and compensated flux accumulation.

Its fields are compared with the CPU oracle (FP64) stepped beside it on the
same input, at the option sets' small sizes; the three-material layout at
its full size (3072x1024, 500 steps), where the oracle would take too long,
is compared with the FP64 form of this library, which the parity tests pin
bitwise to the oracle at that size.  Tolerances (relative L1 of density and
energy <= 1e-5, the precision-form target) are written in each test; the
total mass must stay conserved (MASS_TOL) and the volume fractions in
[0, 1] with unit sum.  Batched mixed problems must equal the same problems
stepped alone bitwise."""
import math
import os

import numpy as np
import pytest

import mahyco_amd
from mahyco_amd import abi, configs, precision
from oracle import refcpu

from gpu_helpers import log
from test_gpu_parity import extension_cases

pytestmark = pytest.mark.gpu

L1_TOL = 1e-5     # relative L1, density and energy (precision-form target)
LINF_TOL = 1e-3   # relative Linf, every reported field
# relative total-mass drift.  The sweeps conserve mass exactly (FP64
# antisymmetric face fluxes, compensated updates); the final remap removes a
# material whose volume fraction falls below the threshold (as the reference
# does), and with FP32 fluxes such a residue is up to ~1e-8 of a cell's volume
# instead of FP64 round-off.
MASS_TOL = 1e-10


def _pair(cfg):
    """The FP64 form of this library and the mixed form on the same input."""
    cfg.precision = 0
    ref = mahyco_amd.Context(cfg)
    cfg.precision = 1
    mix = mahyco_amd.Context(cfg)
    cfg.precision = 0
    return ref, mix


def _oracle_pair(cfg):
    """The CPU oracle (FP64) and the mixed form on the same input."""
    cfg.precision = 0
    orc = refcpu.Oracle(cfg)
    cfg.precision = 1
    mix = mahyco_amd.Context(cfg)
    cfg.precision = 0
    return orc, mix


def _oracle_mass(orc):
    return math.fsum(orc.get("cell_mass"))


def _check(ref, mix, cfg, what, mass0):
    errs = precision.field_errors(mix, ref, cfg.nbmat)
    log(f"[mixed] {what}: " + " ".join(f"{k} {v[0]:.1e}/{v[1]:.1e}" for k, v in errs.items()))
    for k in ("rho", "e"):
        assert errs[k][0] <= L1_TOL, f"{what}: relative L1 of {k} = {errs[k][0]:.3e}"
    for k, (l1, linf) in errs.items():
        assert linf <= LINF_TOL, f"{what}: relative Linf of {k} = {linf:.3e}"
    drift = abs(mix.conservation()[0] - mass0) / mass0
    assert drift <= MASS_TOL, f"{what}: mass drift {drift:.3e}"
    fs = np.stack([mix.get("fracvol", k) for k in range(cfg.nbmat)])
    assert fs.min() >= 0.0 and fs.max() <= 1.0, f"{what}: volume fraction outside [0, 1]"
    assert np.allclose(fs.sum(axis=0), 1.0, rtol=0, atol=1e-12), f"{what}: volume fractions do not sum to 1"


def test_mixed_triple_layout_long_run_against_oracle():
    """The three-material layout (192x64, jittered densities) over 500
    steps, well past the initial dt ramp, against the CPU oracle."""
    cfg = configs.triple_layout(192, 64)
    orc, mix = _oracle_pair(cfg)
    try:
        mass0 = mix.conservation()[0]
        log("[mixed] triple layout 192x64, 500 steps, against the oracle")
        for chunk in (100, 400):
            assert orc.step(chunk) == abi.MHY_OK
            mix.run(chunk)
        ta, tb = orc.time_state(), mix.time_state()
        assert abs(ta.t_n - tb.t_n) <= 1e-5 * ta.t_n
        _check(orc, mix, cfg, "triple layout 500 steps", mass0)
    finally:
        orc.close()
        mix.close()


def test_mixed_triple_layout_full_size():
    """The three-material layout at its full size (3072x1024) over its 500
    steps, against the FP64 form (bitwise the oracle at this size after the
    steps the full-size parity test runs)."""
    cfg = configs.triple_layout()
    ref, mix = _pair(cfg)
    try:
        mass0 = ref.conservation()[0]
        log("[mixed] triple layout 3072x1024, 500 steps, against the FP64 form")
        for chunk in (100, 400):
            ref.run(chunk)
            mix.run(chunk)
        ta, tb = ref.time_state(), mix.time_state()
        assert abs(ta.t_n - tb.t_n) <= 1e-5 * ta.t_n
        _check(ref, mix, cfg, "triple layout 3072x1024, 500 steps", mass0)
    finally:
        ref.close()
        mix.close()


@pytest.mark.parametrize("name", sorted(extension_cases()))
def test_mixed_options(name):
    """Every option set of the parity suite (PenteBorne variants, A1/A2/PB
    dual mass fluxes, mixte flags, limiters, boundary fluxes) over 200
    steps, against the CPU oracle."""
    cfg = extension_cases()[name]
    orc, mix = _oracle_pair(cfg)
    try:
        log(f"[mixed] {name} {cfg.nx}x{cfg.ny} m={cfg.nbmat}, 200 steps, against the oracle")
        assert orc.step(200) == abi.MHY_OK
        mix.run(200)
        # the oracle's total mass (inflow boundaries change the mass)
        _check(orc, mix, cfg, name, _oracle_mass(orc))
    finally:
        orc.close()
        mix.close()


def test_mixed_field_storage_is_fp32():
    """Bulk fields round-trip through FP32 storage; volume fractions and
    masses keep FP64."""
    cfg = configs.riemann(24, 0.45, 1.0, 1.0, 0.125, 0.1, 1.4, 1.67)
    cfg.precision = 1
    ctx = mahyco_amd.Context(cfg)
    try:
        n = cfg.nx * cfg.ny
        rho = 1.0 + np.arange(n) * 1e-9
        ctx.set("density", rho)
        assert np.array_equal(ctx.get("density"), rho.astype(np.float32).astype(np.float64))
        f = np.full(n, 1.0 / 3.0)
        ctx.set("fracvol", f, 0)
        assert np.array_equal(ctx.get("fracvol", 0), f)
    finally:
        ctx.close()


@pytest.mark.parametrize("graphs", [True, False], ids=["graphs", "eager"])
def test_mixed_batch_matches_single_problems(graphs):
    cfgs = configs.ensemble(n=7, side_min=12, side_max=80)
    for c in cfgs:
        c.precision = 1
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
        log(f"[mixed batch] {len(cfgs)} problems, graphs={graphs}")
        for chunk in (6, 6, 3):  # 10 tuning steps, then replays
            batch.run(chunk)
            for s in singles:
                s.step(chunk)
            for i, c in enumerate(cfgs):
                pb, r = batch.problem(i), singles[i]
                for fld in ("density", "energy", "pressure", "velocity_x", "velocity_y"):
                    assert np.array_equal(pb.get(fld), r.get(fld), equal_nan=True), f"problem {i} {fld}"
                for k in range(c.nbmat):
                    assert np.array_equal(pb.get("fracvol", k), r.get("fracvol", k)), f"problem {i} fracvol{k}"
    finally:
        for s in singles:
            s.close()
        batch.close()
