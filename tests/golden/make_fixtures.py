"""Extracts compact copies of the reference's own golden fixtures.

For every reference test directory whose donnees.txt selects the VNR
scheme, copies donnees.txt and packs VnrRemap.pvd plus the first, middle and
last VTP snapshots of reference/ into tests/golden/vnr/<case>.tar.xz
(config C1, SOD_X_Mono_mat_Minmod_vnr_o2, keeps every snapshot).  For the
Eucclhyd scheme, tests/golden/euc/<case>/ gets EucclhydRemap.pvd,
Particules.pvd, the first, middle and last EucclhydRemap snapshots and the
first Particules snapshot.  These are data files the reference's tests hold
(inputs and expected outputs), used so that fixture comparisons run where
the reference tree is absent.

Usage: python tests/golden/make_fixtures.py [/root/reference/tests]
"""
import os
import shutil
import sys
import tarfile

HERE = os.path.dirname(os.path.abspath(__file__))
FULL = {"SOD_X_Mono_mat_Minmod_vnr_o2"}


def scheme_cases(src, scheme):
    for name in sorted(os.listdir(src)):
        d = os.path.join(src, name)
        f = os.path.join(d, "donnees.txt")
        if not os.path.isfile(f):
            continue
        with open(f, encoding="latin-1") as fh:
            if any(line.strip() == scheme for line in fh):
                yield name, d


def vnr_cases(src):
    return scheme_cases(src, "VNR")


def snapshot_key(fn):
    return int(fn.split(".")[1])


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tests"
    out = os.path.join(HERE, "vnr")
    os.makedirs(out, exist_ok=True)
    count = 0
    for name, d in vnr_cases(src):
        ref = os.path.join(d, "reference")
        vtps = sorted((f for f in os.listdir(ref) if f.endswith(".vtp")), key=snapshot_key)
        keep = vtps if name in FULL else sorted({vtps[0], vtps[len(vtps) // 2], vtps[-1]}, key=snapshot_key)
        os.makedirs(os.path.join(out, name), exist_ok=True)
        shutil.copyfile(os.path.join(d, "donnees.txt"), os.path.join(out, name, "donnees.txt"))
        with tarfile.open(os.path.join(out, name, "reference.tar.xz"), "w:xz") as tar:
            for f in ["VnrRemap.pvd"] + keep:
                tar.add(os.path.join(ref, f), arcname=f)
        count += 1
    print(f"{count} VNR fixtures written to {out}")
    out = os.path.join(HERE, "euc")
    os.makedirs(out, exist_ok=True)
    count = 0
    for name, d in scheme_cases(src, "Eucclhyd"):
        ref = os.path.join(d, "reference")
        vtps = sorted((f for f in os.listdir(ref) if f.startswith("EucclhydRemap.") and f.endswith(".vtp")),
                      key=snapshot_key)
        keep = sorted({vtps[0], vtps[len(vtps) // 2], vtps[-1]}, key=snapshot_key)
        keep += ["Particules." + vtps[0].split(".", 1)[1]]
        os.makedirs(os.path.join(out, name), exist_ok=True)
        shutil.copyfile(os.path.join(d, "donnees.txt"), os.path.join(out, name, "donnees.txt"))
        with tarfile.open(os.path.join(out, name, "reference.tar.xz"), "w:xz") as tar:
            for f in ["EucclhydRemap.pvd", "Particules.pvd"] + keep:
                tar.add(os.path.join(ref, f), arcname=f)
        count += 1
    print(f"{count} Eucclhyd fixtures written to {out}")


if __name__ == "__main__":
    main()
