"""Seeded random option sets against the oracle: limiter (all eight),
projection order 1 to 3, PenteBorne with its mixte / complet variants and
pure-cell limiter, the A1 / A2 / PB dual mass fluxes, one to three
materials, and mesh sides that are not multiples of any tile or strip size.
Every cell, material and node field must stay within 1e-10 relative Linf of
the oracle (in practice bitwise) over 15 steps."""
import numpy as np
import pytest

from mahyco_amd import abi, configs

from gpu_helpers import log, run_pair

pytestmark = pytest.mark.gpu
N_CASES = 24
LIMS = sorted(abi.LIM.values())


def random_case(seed):
    rng = np.random.default_rng(seed)
    nx, ny = int(rng.integers(17, 71)), int(rng.integers(9, 41))
    kind = ["riemann", "perturbed", "triple", "point"][seed % 4]
    if kind == "riemann":
        g = [1.4, 1.67]
        c = configs.riemann(nx, float(rng.uniform(0.3, 0.7)), float(rng.uniform(0.1, 1)), float(rng.uniform(0.05, 1)),
                            float(rng.uniform(0.1, 1)), float(rng.uniform(0.05, 1)), g[rng.integers(2)], g[rng.integers(2)])
    elif kind == "perturbed":
        c = configs.riemann_perturbed(n=nx)
    elif kind == "triple":
        c = configs.triple_layout(nx, ny)
    else:
        c = configs.point_explosion(nx)
    c.ny = ny
    c.limiter = int(rng.choice(LIMS))
    c.projection_order = int(rng.choice([1, 2, 2, 3]))
    c.pp = 1 if (c.projection_order <= 2 and rng.uniform() < 0.5) else 0
    if c.pp:
        c.pp_mixte = int(rng.integers(2))
        c.pp_complet = int(rng.integers(2))
        c.limiter_pure = int(rng.choice(LIMS))
        c.methode_flux_masse = int(rng.integers(3))
    else:
        c.pp_mixte = c.pp_complet = 0
        c.methode_flux_masse = int(rng.integers(2))  # the PB dual fluxes come from PenteBorne
    c.abort_on_energy = 0
    desc = (f"{kind} {nx}x{ny} m={c.nbmat} lim={c.limiter} order={c.projection_order} pp={c.pp} "
            f"mixte={c.pp_mixte} complet={c.pp_complet} pure={c.limiter_pure} methode={c.methode_flux_masse}")
    return c, desc


@pytest.mark.parametrize("seed", range(N_CASES))
def test_random_options_step_parity(seed):
    cfg, desc = random_case(seed)
    log(f"[random] seed {seed}: {desc}")
    res = run_pair(cfg, steps=15, checkpoints=(1, 5), tol=1e-10)
    log(f"[random] seed {seed}: worst rel {max(r[1] for r in res):.2e} bitwise={all(r[3] for r in res)}")
