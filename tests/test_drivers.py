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


REFERENCE_MAIN = build.REFERENCE_MAIN
needs_reference = pytest.mark.skipif(not os.path.exists(REFERENCE_MAIN),
                                     reason="the reference sources are not present on this machine")


@needs_reference
def test_reference_main_compiles_with_only_include_and_namespace_changes(tmp_path):
    """The reference's own main.cc (both scheme branches) against
    include/mahyco.hpp: the only lines that change are include lines, the
    using-directive and the namespace of the runtime lifecycle calls."""
    with open(REFERENCE_MAIN) as fh:
        _, changed = build.reference_main_source(fh.read())
    for n, old, new in changed:
        ok = (old.lstrip().startswith("#include") or old.strip() == "using namespace nablalib;"
              or (old.replace("Kokkos::", "mahyco::") == new))
        assert ok, f"line {n} changed beyond include / namespace: {old!r} -> {new!r}"
    exe = build.build_reference_main(str(tmp_path / "refmain"))
    # the reference's command-line errors (main.cc:28-35)
    r = subprocess.run([exe, "a", "b"], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1 and "Fichier de donnees" in r.stderr
    r = subprocess.run([exe, str(tmp_path / "missing.txt")], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1 and "ERREUR" in r.stdout


CONFIG_DUMP = r'''
#include <cstdio>
#include "mahyco.hpp"
using namespace mahyco;
static void dump(const mhy_config& c) {
  std::printf("%d %d %d %.17g %.17g %.17g %.17g %.17g %.17g %d %d", c.test_case, c.nx, c.ny, c.dx, c.dy,
              c.final_time, c.output_time, c.deltat_init, c.deltat_min, c.max_time_iterations, c.nbmat);
  for (int k = 0; k < MHY_NBMATMAX; ++k) std::printf(" %d %.17g", c.eos[k], c.gammap[k]);
  std::printf(" %.17g %d %d %d %d %.17g %d %d %d %d %d %d %d %d %d %d %d %llu", c.gamma, c.schema, c.space_order,
              c.avec_projection, c.projection_conservative, c.threshold, c.projection_order, c.limiter, c.pp,
              c.pp_mixte, c.limiter_pure, c.pp_complet, c.pseudo_centree, c.methode_flux_masse, c.sans_lagrange,
              c.abort_on_energy, c.precision, c.seed);
  for (double v : c.param) std::printf(" %.17g", v);
  std::printf("\n");
}
int main(int argc, char* argv[]) {
  for (int a = 1; a < argc; ++a) {
    mhy_config direct;
    LectureDonneesClass().LectureDonnees(argv[a], &direct);
    schemalagrangelib::SchemaLagrangeClass::SchemaLagrange s;
    optionschemalib::OptionsSchema::Options o;
    cstmeshlib::ConstantesMaillagesClass::ConstantesMaillages cm;
    gesttempslib::GestionTempsClass::GestTemps gt;
    limiteurslib::LimiteursClass::Limiteurs l;
    eoslib::EquationDetat::Eos e;
    castestlib::CasTest::Test t;
    LectureDonneesClass().LectureDonnees(argv[a], &s, &o, &cm, &gt, &l, &e, &t);
    dump(direct);
    dump(optionsToConfig(s.schema, &o, &cm, &gt, &l, &e, &t));
  }
  return 0;
}
'''


def test_option_structs_round_trip_the_reader(tmp_path):
    """The reference-signature reader fills the option structs, and the
    structs give back the configuration the reader produces directly, on
    every VNR test input."""
    src = tmp_path / "dump.cc"
    src.write_text(CONFIG_DUMP)
    exe = tmp_path / "dump"
    if not os.path.exists(build.LIB):
        build.build()
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                        str(exe), "-L", build.LIB_DIR, "-lmahyco_hip", "-Wl,-rpath," + build.LIB_DIR],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    golden = os.path.join(ROOT, "tests", "golden", "vnr")
    inputs = [os.path.join(golden, c, "donnees.txt") for c in sorted(os.listdir(golden))]
    r = subprocess.run([str(exe)] + inputs, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if ln and ln[0].isdigit() or ln.startswith("-")]
    assert len(lines) == 2 * len(inputs)
    for i, path in enumerate(inputs):
        assert lines[2 * i] == lines[2 * i + 1], path
