"""The mahyco executable (C++ drivers over the C ABI, the counterpart of the
reference's main + Vnr + Remap) on the 37 VNR test inputs, run the way the
reference's tests run (tests/launch_test.sh:7-31): `mahyco donnees.txt` in
an empty directory, then the produced output/ is compared with the golden
files byte for byte."""
import os
import subprocess
import tarfile

import pytest

from mahyco_amd import build

from gpu_helpers import log

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "vnr")
CASES = sorted(os.listdir(GOLDEN))


@pytest.mark.parametrize("case", CASES)
def test_driver_output_matches_golden(case, tmp_path):
    assert os.path.exists(build.EXE), "build the driver first (mahyco_amd.build)"
    log(f"[driver] {case}")
    r = subprocess.run([build.EXE, os.path.join(GOLDEN, case, "donnees.txt")], cwd=tmp_path, capture_output=True,
                       text=True, timeout=100)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    out = tmp_path / "output"
    with tarfile.open(os.path.join(GOLDEN, case, "reference.tar.xz")) as tar:
        members = tar.getmembers()
        for m in members:
            want = tar.extractfile(m).read()
            assert (out / m.name).read_bytes() == want, f"{case}: {m.name}"
    produced = {p.name for p in out.iterdir()}
    with open(out / "VnrRemap.pvd") as fh:
        listed = {line.split('file="')[1].split('"')[0] for line in fh if "file=" in line}
    assert listed | {"VnrRemap.pvd"} == produced
    if len(members) == len(produced):  # complete golden set: diff -r equality
        assert {m.name for m in members} == produced
