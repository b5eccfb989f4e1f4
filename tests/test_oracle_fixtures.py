"""The CPU oracle against the reference's own golden outputs.

Each of the 63 test directories of the reference (tests/<case>/donnees.txt
+ reference/*.vtp, compared by the reference with diff -r,
tests/launch_test.sh:24) is replayed through the oracle: the 37 VNR cases
and the 26 Eucclhyd cases (the second caller of the same cell sweeps, whose
fixtures are the reference-held data for PenteBorne-mixte, three materials,
third-order remap and the generalised limiters).  The VTP/PVD text the
oracle writes must be byte-identical to the golden files kept in
tests/golden/vnr and tests/golden/euc (first, middle and last snapshot + the
PVD indexes).  When the reference tree is mounted, every snapshot is
compared.
"""
import os
import subprocess
import tarfile

import pytest

from oracle import refcpu

HERE = os.path.join(os.path.dirname(__file__), "golden")
REF_TESTS = "/root/reference/tests"
SCHEMES = {"vnr": "VnrRemap", "euc": "EucclhydRemap"}
CASES = sorted((scheme, case) for scheme in SCHEMES for case in os.listdir(os.path.join(HERE, scheme)))
IDS = [f"{scheme}-{case}" for scheme, case in CASES]


@pytest.fixture(scope="module", autouse=True)
def built():
    refcpu.build()


_RUNS = {}


@pytest.fixture(scope="module")
def all_runs(tmp_path_factory):
    """Runs the oracle on every case, several single-threaded processes at a
    time (the meshes are tiny), and returns {case: output_dir}."""
    if not _RUNS:
        env = dict(os.environ, OMP_NUM_THREADS="1")
        width = max(1, min(8, os.cpu_count() or 1))
        pending = list(CASES)
        running = []
        while pending or running:
            while pending and len(running) < width:
                case = pending.pop(0)
                out = str(tmp_path_factory.mktemp(case[1]) / "output")
                donnees = os.path.join(HERE, case[0], case[1], "donnees.txt")
                proc = subprocess.Popen([refcpu.RUNNER, donnees, out], stdout=subprocess.PIPE,
                                        stderr=subprocess.STDOUT, env=env)
                running.append((case, out, proc))
            case, out, proc = running.pop(0)
            log = proc.communicate(timeout=600)[0]
            assert proc.returncode == 0, f"{case}: {log!r}"
            _RUNS[case] = out
    return _RUNS


def test_all_63_cases_present():
    assert sum(1 for s, _ in CASES if s == "vnr") == 37
    assert sum(1 for s, _ in CASES if s == "euc") == 26


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_oracle_matches_golden(case, all_runs):
    out = all_runs[case]
    module = SCHEMES[case[0]]
    with tarfile.open(os.path.join(HERE, case[0], case[1], "reference.tar.xz")) as tar:
        members = tar.getmembers()
        assert any(m.name == module + ".pvd" for m in members)
        for m in members:
            want = tar.extractfile(m).read()
            path = os.path.join(out, m.name)
            assert os.path.exists(path), f"{case}: {m.name} not written"
            with open(path, "rb") as fh:
                got = fh.read()
            assert got == want, f"{case}: {m.name} differs from the golden file"
    # every snapshot listed in the PVD indexes exists, and nothing else was written
    listed = set()
    for pvd in [f for f in os.listdir(out) if f.endswith(".pvd")]:
        with open(os.path.join(out, pvd)) as fh:
            listed |= {line.split('file="')[1].split('"')[0] for line in fh if "file=" in line}
    written = {f for f in os.listdir(out) if f.endswith(".vtp")}
    assert listed == written


@pytest.mark.skipif(not os.path.isdir(REF_TESTS), reason="reference tree not mounted")
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_oracle_matches_full_reference_tree(case, all_runs):
    out = all_runs[case]
    ref = os.path.join(REF_TESTS, case[1], "reference")
    assert sorted(os.listdir(out)) == sorted(os.listdir(ref))
    for f in os.listdir(ref):
        with open(os.path.join(ref, f), "rb") as a, open(os.path.join(out, f), "rb") as b:
            assert a.read() == b.read(), f"{case}: {f}"
