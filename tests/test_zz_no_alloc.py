"""mhy_conservation_noalloc through the C ABI: the conservation sums with the
block partials in the library's static per-GPU scratch array instead of a
buffer allocated per call.  Its results must equal mhy_conservation's
bitwise, on two calls in a row, the second on a smaller mesh (fewer blocks:
the first call's partials are still in the array and must not be read)."""
import ctypes

import numpy as np
import pytest

import mahyco_amd
from mahyco_amd import configs

from gpu_helpers import log

pytestmark = pytest.mark.gpu


def _sums(ctx, entry):
    out = (ctypes.c_double * 5)()
    st = getattr(mahyco_amd.library(), entry)(ctx.h, out)
    assert st == 0, f"{entry}: status {st}: {mahyco_amd.library().mhy_last_error(ctx.h).decode()}"
    return np.array(out[: 2 + ctx.cfg.nbmat])


def test_conservation_noalloc_equals_allocating_entry_point():
    big = configs.triple_layout(768, 256)     # 196608 cells: 256 blocks of partials
    small = configs.riemann(20, 0.45, 1.0, 1.0, 0.125, 0.1, 1.4, 1.67)  # 400 cells: 2 blocks
    for cfg in (big, small):
        ctx = mahyco_amd.Context(cfg)
        try:
            ctx.run(3)
            want = _sums(ctx, "mhy_conservation")
            got = _sums(ctx, "mhy_conservation_noalloc")
            log(f"[noalloc] {cfg.nx}x{cfg.ny} m={cfg.nbmat}: {got.tolist()}")
            assert np.array_equal(got, want), (got, want)
            assert np.all(np.isfinite(got)) and got[0] > 0
        finally:
            ctx.close()
