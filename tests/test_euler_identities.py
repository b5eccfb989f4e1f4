"""Floating-point identities the kernels rely on to skip work on the fixed
Euler grid while staying bitwise equal to the reference arithmetic.

The node kernel forms the corner normals of the Euler grid as the plain sum
of the two half-edge vectors instead of computeLpcNpc's length x (vector /
length) (includes/Freefunctions.h:25-58, PhaseLagrange.cc:114-136).  That is
exact because every half edge is axis-aligned: one component is +0, the other
a, with sqrt(fl(a*a)) = |a| in binary floating point, a/|a| = +-1 and
(+-1)*|a| = a.  Checked here in IEEE binary64 (numpy) on random values and on
every node of the meshes of the workloads and fixtures."""
import numpy as np
import pytest


def _lpc(xp, xpp, xpm):
    # computeLpcNpc as the reference writes it (dot product accumulated from 0.0)
    n1 = (0.5 * (xpp[1] - xp[1]), 0.5 * (xp[0] - xpp[0]))
    l1 = np.sqrt(0.0 + n1[0] * n1[0] + n1[1] * n1[1])
    n2 = (0.5 * (xp[1] - xpm[1]), 0.5 * (xpm[0] - xp[0]))
    l2 = np.sqrt(0.0 + n2[0] * n2[0] + n2[1] * n2[1])
    return ((n1[0] / l1) * l1 + (n2[0] / l2) * l2, (n1[1] / l1) * l1 + (n2[1] / l2) * l2)


def _axis(xp, xpp, xpm):
    # the kernel's form on the Euler grid (lpcNpcAxis)
    return (0.5 * (xpp[1] - xp[1]) + 0.5 * (xp[1] - xpm[1]), 0.5 * (xp[0] - xpp[0]) + 0.5 * (xpm[0] - xp[0]))


def _same(a, b):
    return (a == b) & (np.signbit(a) == np.signbit(b))


def test_sqrt_of_square_is_abs():
    rng = np.random.default_rng(7)
    a = rng.standard_normal(2_000_000) * np.exp2(rng.integers(-200, 200, 2_000_000).astype(np.float64))
    assert np.array_equal(np.sqrt(a * a), np.abs(a))


@pytest.mark.parametrize("nx,ny,lx,ly", [(4096, 4096, 1.0, 1.0), (2048, 2048, 1.0, 1.0), (3072, 1024, 3.0, 1.0),
                                         (200, 5, 1.0, 0.1), (5, 200, 0.1, 1.0), (100, 10, 1.0, 0.1),
                                         (1024, 1024, 1.0, 1.0), (977, 613, 1.3, 0.7)])
def test_axis_corner_normal_equals_lpcnpc(nx, ny, lx, ly):
    dx, dy = lx / nx, ly / ny
    I, J = np.meshgrid(np.arange(nx + 1, dtype=np.float64), np.arange(ny + 1, dtype=np.float64))

    def X(jj, ii):  # CartesianMesh2DGenerator.cc:52-55
        return (dx * ii, dy * jj)

    # the four corners of a node: (+ neighbour, - neighbour) offsets as in the node kernel
    for (pi, pj, mi, mj) in [(1, 0, 0, 1), (0, 1, -1, 0), (0, -1, 1, 0), (-1, 0, 0, -1)]:
        xp, xpp, xpm = X(J, I), X(J + pj, I + pi), X(J + mj, I + mi)
        r, s = _lpc(xp, xpp, xpm), _axis(xp, xpp, xpm)
        assert _same(r[0], s[0]).all() and _same(r[1], s[1]).all(), (pi, pj, mi, mj)
