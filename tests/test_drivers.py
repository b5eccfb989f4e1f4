"""The C++ drivers of include/mahyco.hpp and the mahyco executable, without a
GPU: the header compiles on its own, and the executable keeps the
reference's command-line error behaviour (main.cc:28-33: no input file ->
message and exit status 1; an unreadable input -> message and exit 1)."""
import os
import subprocess

import pytest

from mahyco_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(build.EXE):
        if not os.path.exists(build.LIB):
            build.build()
        else:
            build.build_driver()
    return build.EXE


def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "use.cc"
    src.write_text('#include "mahyco.hpp"\nint main() { mahyco::LectureDonneesClass l; (void)l; return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_no_argument_exits_1(exe, tmp_path):
    r = subprocess.run([exe], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1
    assert "Fichier de donnees" in r.stderr


def test_unreadable_input_exits_1(exe, tmp_path):
    r = subprocess.run([exe, str(tmp_path / "missing.txt")], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1
    assert "ERREUR" in r.stdout
    assert not (tmp_path / "output").exists()
