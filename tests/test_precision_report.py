"""Host-side pieces of the mixed-form error report (no GPU): the relative
error definitions and the algorithmic bytes per cell-update of both forms."""
import numpy as np
import pytest

import bench
from mahyco_amd import precision


def test_relative_errors_definitions():
    b = np.array([1.0, -2.0, 4.0])
    a = b + np.array([0.1, 0.0, -0.2])
    l1, linf = precision.relative_errors(a, b)
    assert l1 == pytest.approx(0.3 / 7.0, rel=1e-15)
    assert linf == pytest.approx(0.2 / 4.0, rel=1e-15)
    assert precision.relative_errors(b, b) == (0.0, 0.0)


def test_velocity_errors_use_the_speed():
    # a component that is zero in the reference is measured against the speed
    ux = np.array([3.0, 0.0])
    uy = np.array([0.0, 4.0])
    speed = np.hypot(ux, uy)
    l1, linf = precision.relative_errors(uy + 1e-3, uy, speed)
    assert l1 == pytest.approx(2e-3 / 7.0, rel=1e-12)
    assert linf == pytest.approx(1e-3 / 4.0, rel=1e-12)


def test_report_lists_every_material_fraction():
    labels = [x[0] for x in precision.report_fields(3)]
    assert labels == ["rho", "e", "p", "ux", "uy", "f_1", "f_2", "f_3"]


@pytest.mark.parametrize("m,fp64,mixed", [(1, 264, 156), (2, 376, 228), (3, 488, 300)])
def test_algorithmic_bytes(m, fp64, mixed):
    # persistent words per cell-update: 17 + 14 m words of 8 B in FP64 (+16 B
    # of flags); in the mixed form f_k and M_k (read and written: 4 m words)
    # stay 8 B and the other 17 + 10 m words take 4 B
    assert bench.algorithmic_bytes_per_cell(m, 0) == fp64
    assert bench.algorithmic_bytes_per_cell(m, 1) == mixed
    assert bench.algorithmic_bytes_per_cell(m, 1) == 4 * (17 + 10 * m) + 8 * 4 * m + 16
