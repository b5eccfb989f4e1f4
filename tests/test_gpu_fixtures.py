"""The HIP path through the drop-in time loop (mhy_simulate, the
counterpart of Vnr::simulate) on the reference's 37 VNR test inputs: the
VTP/PVD text it writes must be byte-identical to the golden outputs, like
the reference's own tests (diff -r, tests/launch_test.sh:24)."""
import os
import tarfile

import pytest

import mahyco_amd

from gpu_helpers import log

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "vnr")
CASES = sorted(os.listdir(GOLDEN))


@pytest.mark.parametrize("case", CASES)
def test_gpu_simulate_matches_golden(case, tmp_path):
    cfg = mahyco_amd.read_donnees(os.path.join(GOLDEN, case, "donnees.txt"))
    log(f"[fixture] {case}")
    out = str(tmp_path / "output")
    ctx = mahyco_amd.Context(cfg)
    try:
        ctx.simulate(out)
    finally:
        ctx.close()
    with tarfile.open(os.path.join(GOLDEN, case, "reference.tar.xz")) as tar:
        for m in tar.getmembers():
            want = tar.extractfile(m).read()
            with open(os.path.join(out, m.name), "rb") as fh:
                assert fh.read() == want, f"{case}: {m.name}"
    with open(os.path.join(out, "VnrRemap.pvd")) as fh:
        listed = {line.split('file="')[1].split('"')[0] for line in fh if "file=" in line}
    assert listed == {f for f in os.listdir(out) if f.endswith(".vtp")}
