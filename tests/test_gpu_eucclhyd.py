"""The Eucclhyd scheme on the HIP path (eucclhyd.hip and the cell sweeps on
the Eucclhyd remap-vector layout).

- The reference's 26 Eucclhyd test inputs through the drop-in time loop
  (mhy_simulate, the counterpart of Eucclhyd::simulate) and through the
  mahyco executable and the reference's main.cc built against
  include/mahyco.hpp: the VTP/PVD text must be byte-identical to the golden
  outputs (tests/launch_test.sh:24).
- Field by field against the CPU oracle, bitwise, on meshes and option sets
  the fixtures do not cover: larger meshes, PenteBorne-mixte and -complet,
  third order, the generalised limiters, first space order, a conservative
  projection (its kinetic-energy component) and the Lagrange-only mode.
"""
import os
import subprocess
import tarfile

import numpy as np
import pytest

import mahyco_amd
from mahyco_amd import abi, build
from oracle import refcpu

from gpu_helpers import log, rel_linf

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "euc")
CASES = sorted(os.listdir(GOLDEN))

CELL = ["density", "pressure", "energy", "sound_speed", "cell_mass", "cell_velocity_x", "cell_velocity_y",
        "cell_out_vx", "cell_out_vy", "mixte", "pure"]
MAT = ["fracvol", "mass_fraction", "density_env", "energy_env", "pressure_env", "mass_env", "sound_speed_env"]
NODE = ["velocity_x", "velocity_y", "coord_x", "coord_y"]


def _check_golden(case, out):
    with tarfile.open(os.path.join(GOLDEN, case, "reference.tar.xz")) as tar:
        members = tar.getmembers()
        for m in members:
            want = tar.extractfile(m).read()
            with open(os.path.join(out, m.name), "rb") as fh:
                assert fh.read() == want, f"{case}: {m.name}"
    listed = set()
    for pvd in ("EucclhydRemap.pvd", "Particules.pvd"):
        with open(os.path.join(out, pvd)) as fh:
            listed |= {line.split('file="')[1].split('"')[0] for line in fh if "file=" in line}
    assert listed == {f for f in os.listdir(out) if f.endswith(".vtp")}
    return members


@pytest.mark.parametrize("case", CASES)
def test_gpu_simulate_matches_golden(case, tmp_path):
    cfg = mahyco_amd.read_donnees(os.path.join(GOLDEN, case, "donnees.txt"))
    assert cfg.schema == abi.SCHEMA_EUCCLHYD
    log(f"[euc fixture] {case}")
    out = str(tmp_path / "output")
    ctx = mahyco_amd.Context(cfg)
    try:
        ctx.simulate(out)
    finally:
        ctx.close()
    _check_golden(case, out)


EXES = {"mahyco": build.EXE, "reference_main": build.REFMAIN_EXE}


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("which", sorted(EXES))
def test_driver_output_matches_golden(which, case, tmp_path):
    exe = EXES[which]
    if which == "reference_main" and not os.path.exists(exe):
        pytest.skip("the reference's main.cc was not available where this tree was built")
    assert os.path.exists(exe), "build the driver first (mahyco_amd.build)"
    log(f"[euc driver] {which} {case}")
    r = subprocess.run([exe, os.path.join(GOLDEN, case, "donnees.txt")], cwd=tmp_path, capture_output=True,
                       text=True, timeout=100)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    _check_golden(case, str(tmp_path / "output"))


def _fields(cfg):
    names = [(n, 0) for n in CELL] + [(n, 0) for n in NODE]
    names += [(n, k) for k in range(cfg.nbmat) for n in MAT]
    return names


def _from(case, **kw):
    cfg = mahyco_amd.read_donnees(os.path.join(GOLDEN, case, "donnees.txt"))
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def _parity_cases():
    lim = abi.LIM
    return {
        "bisod_x_256x8_pp_mixte": (_from("SOD_X_bi_mat_Superbee_PenteBorne_mix", nx=256, ny=8, dx=1.0 / 256), 200),
        "bisod_y_8x256_pp_complet": (_from("SOD_Y_bi_mat_Superbee_PenteBorne_pure_complet", nx=8, ny=256,
                                           dy=1.0 / 256), 200),
        "sedov_96_pp_mixte": (_from("SEDOV_Mono_Mat_Superbee_PenteBorne_mixte", nx=96, ny=96, dx=1.2 / 96,
                                    dy=1.2 / 96), 120),
        "sedov_64_ultrabeeG": (_from("SEDOV_Mono_Mat_MinModG_simple", nx=64, ny=64, dx=1.2 / 64, dy=1.2 / 64,
                                     limiter=lim["ultrabeeG"]), 120),
        "triple_140x60_order3": (_from("POINT_TRIPLE_proj_ordre_3", nx=140, ny=60, dx=0.0005, dy=0.0005), 150),
        "triple_140x60_pp_mixte": (_from("POINT_TRIPLE_SuperBee_PenteBorne_mixte", nx=140, ny=60, dx=0.0005,
                                         dy=0.0005), 150),
        "triple_70x30_pp_complet_mixte": (_from("POINT_TRIPLE_SuperBee_PenteBorne_mixte", pp_complet=1), 200),
        "bisod_x_conservative": (_from("SOD_X_bi_mat_MinMod_simple", nx=256, ny=8, dx=1.0 / 256,
                                       projection_conservative=1), 200),
        "bisod_x_pp_conservative": (_from("SOD_X_bi_mat_Superbee_PenteBorne_pure", projection_conservative=1),
                                    200),
        "bisod_x_space1_order1": (_from("SOD_X_bi_mat_MinMod_simple", space_order=1, projection_order=1), 200),
        "bisod_x_lagrange_only_256": (_from("SOD_X_bi_mat_Lagrange_Euc", nx=256, ny=8, dx=1.0 / 256), 200),
        "sod_y_mono_vanleer_order2": (_from("SOD_Y_Mono_mat_proj_ordre_3", projection_order=2,
                                            limiter=lim["vanLeer"]), 200),
    }


@pytest.mark.parametrize("name", sorted(_parity_cases()))
def test_fields_bitwise_equal_to_oracle(name):
    cfg, steps = _parity_cases()[name]
    log(f"[euc parity] {name} {cfg.nx}x{cfg.ny} m={cfg.nbmat} steps={steps}")
    gpu = mahyco_amd.Context(cfg)
    orc = refcpu.Oracle(cfg)
    try:
        done = 0
        for target in (1, 7, steps // 2, steps):
            n = target - done
            gpu.step(n)
            assert orc.step(n) == abi.MHY_OK
            done = target
            tg, to = gpu.time_state(), orc.time_state()
            assert (tg.t_n, tg.deltat_n) == (to.t_n, to.deltat_n), (target, tg.t_n, to.t_n)
            diff = []
            for fname, k in _fields(cfg):
                a, b = gpu.get(fname, k), orc.get(fname, k)
                if not np.array_equal(a, b, equal_nan=True):
                    diff.append(f"{fname}[{k}] rel {rel_linf(a, b):.2e}")
            assert not diff, f"after {target} steps: {diff}"
    finally:
        gpu.close()
        orc.close()


def test_bisod_1024_pp_mixte_properties():
    """Full-size run (two-material Sod, 1024^2 cells on the unit square,
    PenteBorne-mixte, 300 steps) checked through what the scheme preserves:
    each material's mass (walls all round) to 1e-8 relative, fractions in
    [0, 1] summing to 1, finite positive densities and energies.  The remap's
    final stage rescales each cell's material volumes to the cell volume and
    renormalises the fractions (eucclhyd_remap/PhaseRemapFinal.cc:17-232),
    so material masses are kept only to rounding drift: the oracle drifts by
    2e-9 relative on this case at 256^2 over 300 steps.  (The reference fixes the first
    time step of its cases, 1e-4 here, about the CFL step of this mesh; a
    Sedov case at this size starts far above it and stops on the NaN check,
    in the reference as here.)"""
    cfg = _from("SOD_X_bi_mat_Superbee_PenteBorne_mix", nx=1024, ny=1024, dx=1.0 / 1024, dy=1.0 / 1024)
    log("[euc full size] bi-material Sod 1024^2")
    ctx = mahyco_amd.Context(cfg)
    try:
        vol = cfg.dx * cfg.dy

        def masses():
            return [np.sum(ctx.get("fracvol", k) * ctx.get("density_env", k)) * vol for k in range(cfg.nbmat)]

        m0 = masses()
        ctx.run(300)
        m1 = masses()
        rho, e = ctx.get("density"), ctx.get("energy")
        assert np.all(np.isfinite(rho)) and np.all(np.isfinite(e))
        assert np.all(rho > 0) and np.all(e > 0)
        fs = [ctx.get("fracvol", k) for k in range(cfg.nbmat)]
        assert all(np.all((f >= 0) & (f <= 1)) for f in fs)
        assert np.max(np.abs(sum(fs) - 1.0)) <= 1e-12
        for a, b in zip(m0, m1):
            assert abs(b - a) <= 1e-8 * a, (m0, m1)
        ts = ctx.time_state()
        assert ts.iteration == 300 and ts.t_n > 0
    finally:
        ctx.close()
