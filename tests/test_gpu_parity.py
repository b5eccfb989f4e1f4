"""Parity of the HIP path with the CPU oracle (which reproduces the
reference's golden outputs): same configuration, same seeded initial state,
every field compared after a number of steps.  The bar is 1e-10 relative
L-infinity (the FP64 kernels are built to be bitwise equal; the tests also
report whether they are)."""
import os

import pytest

import mahyco_amd
from mahyco_amd import abi, configs
from oracle import refcpu

from gpu_helpers import log, run_pair

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "vnr")
CASES = sorted(os.listdir(GOLDEN))
TOL = 1e-10


@pytest.fixture(scope="module", autouse=True)
def libs():
    refcpu.lib()
    mahyco_amd.library()


@pytest.mark.parametrize("case", CASES)
def test_fixture_configs_step_parity(case):
    cfg = refcpu.read_donnees(os.path.join(GOLDEN, case, "donnees.txt"))
    log(f"[parity] {case} {cfg.nx}x{cfg.ny} nbmat={cfg.nbmat}")
    res = run_pair(cfg, steps=60, checkpoints=(1, 2, 10, 30), tol=TOL)
    log(f"[parity] {case}: worst rel {max(r[1] for r in res):.2e} bitwise={all(r[3] for r in res)}")


def _small(cfg, nx, ny=None):
    ny = ny or nx
    cfg.dx = cfg.dx * cfg.nx / nx
    cfg.dy = cfg.dy * cfg.ny / ny
    cfg.nx, cfg.ny = nx, ny
    return cfg


def extension_cases():
    out = {}
    out["C2_pp_superbee"] = _small(configs.riemann_perturbed(), 64)
    c = _small(configs.riemann_perturbed(mixte=1), 64)
    out["C2_pp_mixte_arithmeticG"] = c
    c = _small(configs.riemann_perturbed(), 48)
    c.pp_complet = 1
    out["C2_pp_complet"] = c
    c = _small(configs.riemann_perturbed(), 48)
    c.methode_flux_masse = 2
    out["C2_pp_PB"] = c
    c = _small(configs.riemann_perturbed(), 48)
    c.methode_flux_masse = 1
    out["C2_pp_A2"] = c
    c = _small(configs.riemann_perturbed(mixte=1), 48)
    c.pp_complet = 1
    c.limiter_pure = abi.LIM["minmodG"]
    out["C2_pp_mixte_complet_minmodG"] = c
    out["C3_point_explosion"] = _small(configs.point_explosion(), 64)
    out["C4_triple_layout"] = _small(configs.triple_layout(), 96, 32)
    c = _small(configs.triple_layout(), 60, 20)
    c.pp = 1
    c.limiter = abi.LIM["superBee"]
    c.limiter_pure = abi.LIM["vanLeerG"]
    c.pp_mixte = 1
    out["C4_pp_mixte_vanLeerG"] = c
    out["riemann_member"] = configs.riemann(40, 0.43, 0.7, 0.6, 0.2, 0.1, 1.4, 1.67)
    c = configs.riemann(40, 0.5, 1.0, 1.0, 0.125, 0.1, 1.4, 1.4)
    c.limiter = abi.LIM["vanLeer"]
    out["riemann_vanleer"] = c
    c = configs.riemann(40, 0.5, 1.0, 1.0, 0.125, 0.1, 1.4, 1.4)
    c.projection_order = 1
    out["riemann_order1"] = c
    c = configs.riemann(32, 0.5, 1.0, 1.0, 0.125, 0.1, 1.4, 1.4)
    c.limiter = abi.LIM["superBeeG"]
    out["riemann_superBeeG_main"] = c
    # third-order remap (PhaseRemap1.cc:365-389): two materials along x,
    # one material in both directions, three materials on a 2-D layout
    c = configs.riemann(40, 0.5, 1.0, 1.0, 0.125, 0.1, 1.4, 1.67)
    c.projection_order = 3
    out["riemann_order3"] = c
    c = _small(configs.point_explosion(), 32)
    c.projection_order = 3
    out["C3_point_explosion_order3"] = c
    c = _small(configs.triple_layout(), 48, 16)
    c.projection_order = 3
    out["C4_triple_layout_order3"] = c
    c = configs.riemann(40, 0.43, 0.7, 0.6, 0.2, 0.1, 1.4, 1.67)
    c.projection_order, c.methode_flux_masse = 3, 1  # A2: volume fluxes to the dual remap
    out["riemann_order3_A2"] = c
    c = configs.riemann(32, 0.5, 1.0, 1.0, 0.125, 0.1, 1.4, 1.4)
    c.pseudo_centree = 1
    c.abort_on_energy = 1
    out["riemann_pseudo_centree"] = c
    # reference UnitTestCase: imposed velocities and an inflow flux through the left edge
    c = configs.default_config()
    c.test_case, c.schema, c.nbmat = abi.CASE["UnitTestCase"], abi.SCHEMA_VNR, 1
    c.nx, c.ny, c.dx, c.dy = 40, 4, 0.025, 0.025
    c.limiter, c.pseudo_centree, c.abort_on_energy = abi.LIM["minmod"], 0, 0
    out["unit_test_boundary_flux"] = c
    c = configs.default_config()
    c.test_case, c.schema, c.nbmat = abi.CASE["UnitTestCase"], abi.SCHEMA_VNR, 1
    c.nx, c.ny, c.dx, c.dy = 40, 4, 0.025, 0.025
    c.limiter, c.pseudo_centree, c.abort_on_energy = abi.LIM["minmod"], 0, 0
    c.projection_order = 3
    out["unit_test_boundary_flux_order3"] = c
    return out


EXT = extension_cases()


@pytest.mark.parametrize("name", sorted(EXT))
def test_extension_configs_step_parity(name):
    cfg = EXT[name]
    log(f"[parity] {name} {cfg.nx}x{cfg.ny} nbmat={cfg.nbmat} pp={cfg.pp}")
    res = run_pair(cfg, steps=40, checkpoints=(1, 5, 20), tol=TOL)
    log(f"[parity] {name}: worst rel {max(r[1] for r in res):.2e} bitwise={all(r[3] for r in res)}")


def error_cases():
    out = {}
    # a void material next to a perfect gas: the void EOS zeroes the mixture
    # energy (PhaseLagrange.cc:421-430), the remap then divides 0 by 0 and the
    # reference stops on the NaN check (PhaseRemapFinal.cc:174-196)
    c = configs.riemann(32, 0.5, 1.0, 1.0, 0.125, 0.1, 1.4, 1.4)
    c.eos[1] = abi.EOS_VOID
    out["void_eos_nan"] = (c, abi.MHY_ERR_NAN)
    # |e_k| > 5 with the reference's abort switched on (PhaseLagrange.cc:358-369)
    c = _small(configs.point_explosion(), 32)
    c.abort_on_energy = 1
    out["energy_abort"] = (c, abi.MHY_ERR_ENERGY)
    return out


ERR = error_cases()


def rejected_cases():
    out = {}
    c = configs.riemann(16, 0.5, 1.0, 1.0, 0.125, 0.1, 1.4, 1.4)
    c.projection_order, c.pp = 3, 1
    out["order3_pentborne"] = c
    c = configs.riemann(16, 0.5, 1.0, 1.0, 0.125, 0.1, 1.4, 1.4)
    c.projection_order = 4
    out["order4"] = c
    return out


REJ = rejected_cases()


@pytest.mark.parametrize("name", sorted(REJ))
def test_unsupported_options_are_rejected(name):
    with pytest.raises(mahyco_amd.MahycoError) as e:
        mahyco_amd.Context(REJ[name])
    assert f"({abi.MHY_ERR_ARG})" in str(e.value)


@pytest.mark.parametrize("name", sorted(ERR))
def test_error_status_parity(name):
    """Both paths stop at the same step with the same status."""
    cfg, want = ERR[name]
    gpu = mahyco_amd.Context(cfg)
    orc = refcpu.Oracle(cfg)
    try:
        for step in range(1, 11):
            so = orc.step(1)
            try:
                gpu.step(1)
                sg = abi.MHY_OK
            except mahyco_amd.MahycoError as e:
                sg = int(str(e).split("(")[1].split(")")[0])
            assert sg == so, f"step {step}: gpu {sg} oracle {so}"
            if so != abi.MHY_OK:
                break
        assert so == want
    finally:
        gpu.close()
        orc.close()
