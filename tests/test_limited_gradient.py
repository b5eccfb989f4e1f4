"""The sweeps' limited gradient (sweep_math.hpp limitedGradient, the
reference's computeAndLimitGradPhi, remap/UtilesRemap-Impl.h:99-126) built
for the host and checked bit for bit:

* minmod forms one quotient instead of the reference's two (the one whose
  numerator has the smaller magnitude); the result must be bitwise that of
  the oracle's two-quotient form, on special values (signed zeros, the
  1e-12 threshold of tdivide and its neighbours, subnormals, equal and
  one-ulp-apart magnitudes, huge values, infinities, NaN) and on random
  pairs across forty decades;
* every classic limiter, in double and in float, against the two-quotient
  form built from the same header's limiter and tdivide.
"""
import math
import os
import struct
import subprocess

import numpy as np
import pytest

from mahyco_amd import abi, build
from oracle import refcpu

HERE = os.path.dirname(os.path.abspath(__file__))
LIMS = ["minmod", "superBee", "vanLeer"]


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("lg") / "limited_gradient")
    cmd = [build.hipcc(), "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-ffp-contract=off",
           "-I", os.path.join(os.path.dirname(HERE), "mahyco_amd", "csrc"),
           os.path.join(HERE, "host", "limited_gradient.cc"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def _pairs():
    t = 1.0e-12
    special = [0.0, -0.0, t, -t, math.nextafter(t, 0.0), math.nextafter(t, 1.0), -math.nextafter(t, 0.0), 1e-13,
               5e-324, -5e-324, 2.2250738585072014e-308, 1e-300, 0.5, 1.0, -1.0, math.nextafter(1.0, 0.0),
               math.nextafter(1.0, 2.0), 2.0, -3.0, 1e300, -1e300, 1.7976931348623157e308, math.inf, -math.inf,
               math.nan]
    pairs = [(a, b) for a in special for b in special]
    rng = np.random.default_rng(20261203)
    n = 20000
    mag = 10.0 ** rng.uniform(-20, 20, (n, 2))
    sgn = rng.choice([-1.0, 1.0], (n, 2))
    v = mag * sgn
    pairs += [tuple(x) for x in v]
    # equal and nearly equal magnitudes, either sign
    for a in v[:3000, 0]:
        k = int(rng.integers(-3, 4))
        b = a
        for _ in range(abs(k)):
            b = math.nextafter(b, math.inf if k > 0 else -math.inf)
        pairs.append((a, float(rng.choice([-1.0, 1.0])) * b))
    return pairs


def _run(exe, pairs):
    payload = struct.pack("<q", len(pairs)) + np.asarray(pairs, dtype="<f8").tobytes()
    out = subprocess.run([exe], input=payload, capture_output=True, check=True).stdout
    rec = np.dtype([("d", "<f8", 2), ("f", "<f4", 2)])
    return np.frombuffer(out, dtype=rec).reshape(len(LIMS), len(pairs))


def _same_bits(a, b):
    """Bitwise equality, any NaN equal to any NaN."""
    ia = a.view(np.int64 if a.dtype == np.float64 else np.int32)
    ib = b.view(np.int64 if b.dtype == np.float64 else np.int32)
    return (ia == ib) | (np.isnan(a) & np.isnan(b))


def test_every_classic_limiter_equals_the_two_quotient_form(driver):
    pairs = _pairs()
    res = _run(driver, pairs)
    for li, lim in enumerate(LIMS):
        for form in ("d", "f"):
            got, want = res[li][form][:, 0], res[li][form][:, 1]
            bad = np.nonzero(~_same_bits(got, want))[0]
            assert bad.size == 0, (lim, form, [(pairs[i], got[i], want[i]) for i in bad[:5]])


def test_minmod_equals_the_oracle_bitwise(driver):
    pairs = _pairs()
    res = _run(driver, pairs)
    got = res[LIMS.index("minmod")]["d"][:, 0]
    orc = refcpu.Oracle(refcpu_config())
    try:
        want = np.array([orc.probe_gradient(abi.LIM["minmod"], gp, gm, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0)
                         for gp, gm in pairs])
    finally:
        orc.close()
    bad = np.nonzero(~_same_bits(got, want))[0]
    assert bad.size == 0, [(pairs[i], got[i], want[i]) for i in bad[:5]]
    # the special values do reach both the threshold and the sign cases
    assert np.any(got == 0.0) and np.any(np.isnan(got)) and np.any(np.signbit(got) & (got == 0.0))


def refcpu_config():
    from mahyco_amd import configs
    return configs.riemann(8, 0.5, 1.0, 1.0, 0.125, 0.1, 1.4, 1.4)
