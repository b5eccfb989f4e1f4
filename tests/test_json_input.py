"""The JSON input (mhy_read_json, no GPU needed): the workloads' JSON
documents written by mahyco_amd.configs.to_json read back field for field
(bitwise for every double), a donnees.txt base file with overrides, named
options, the ensemble form, and the errors (unknown key, malformed text,
capacity)."""
import ctypes
import json
import os

import pytest

import mahyco_amd
from mahyco_amd import abi, configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOD = os.path.join(ROOT, "tests", "golden", "vnr", "SOD_X_Mono_mat_Minmod_vnr_o2", "donnees.txt")


def _same(a, b):
    return bytes(a) == bytes(b)


@pytest.mark.parametrize("name", sorted(configs.WORKLOADS))
def test_workload_json_round_trip(name, tmp_path):
    cfgs = configs.WORKLOADS[name]()
    p = tmp_path / f"{name}.json"
    p.write_text(configs.to_json(cfgs))
    got = mahyco_amd.read_json(str(p))
    assert len(got) == len(cfgs)
    for i, (a, b) in enumerate(zip(got, cfgs)):
        assert _same(a, b), f"{name} problem {i}: {configs.as_json_dict(a)} != {configs.as_json_dict(b)}"


def test_base_file_with_overrides(tmp_path):
    base = mahyco_amd.read_donnees(SOD)
    p = tmp_path / "sod.json"
    p.write_text(json.dumps({"donnees": os.path.relpath(SOD, tmp_path), "nx": 400, "precision": "mixed",
                             "gammap": [1.67], "limiter": "superBee", "abort_on_energy": False}))
    (c,) = mahyco_amd.read_json(str(p))
    assert (c.nx, c.precision, c.gammap[0], c.limiter, c.abort_on_energy) == (400, 1, 1.67, abi.LIM["superBee"], 0)
    for f in ("test_case", "ny", "dx", "dy", "final_time", "output_time", "nbmat", "schema", "pseudo_centree",
              "methode_flux_masse", "projection_order"):
        assert getattr(c, f) == getattr(base, f), f


def test_named_options_and_case_defaults(tmp_path):
    p = tmp_path / "c.json"
    p.write_text(json.dumps({"test_case": "TripleLayout", "nx": 60, "ny": 20, "dx": 0.05, "dy": 0.05,
                             "pp": "oui", "pp_mixte": 1, "limiter_pure": "vanLeerG", "methode_flux_masse": "PB",
                             "seed": 18446744073709551615, "param": [1e-3]}))
    (c,) = mahyco_amd.read_json(str(p))
    assert c.test_case == abi.CASE["TripleLayout"] and c.nbmat == 3 and c.schema == abi.SCHEMA_VNR
    assert (c.pp, c.pp_mixte, c.limiter_pure, c.methode_flux_masse) == (1, 1, abi.LIM["vanLeerG"], 2)
    assert c.seed == 2 ** 64 - 1 and c.param[0] == 1e-3 and c.param[1] == 0.0


def test_ensemble_problems_inherit_the_top_level(tmp_path):
    p = tmp_path / "e.json"
    p.write_text(json.dumps({"test_case": "Riemann", "limiter": "vanLeer", "max_time_iterations": 5,
                             "problems": [{"nx": 16, "ny": 16, "dx": 1 / 16, "dy": 1 / 16},
                                          {"nx": 32, "ny": 32, "dx": 1 / 32, "dy": 1 / 32, "limiter": "minmod"}]}))
    a, b = mahyco_amd.read_json(str(p))
    assert (a.nx, b.nx) == (16, 32) and (a.nbmat, b.nbmat) == (2, 2)
    assert (a.limiter, b.limiter) == (abi.LIM["vanLeer"], abi.LIM["minmod"])
    assert a.max_time_iterations == b.max_time_iterations == 5


@pytest.mark.parametrize("text", ['{"nx": 10, "nxx": 3}', '{"nx": 10', '{"nx": 1.5}', '{"limiter": "fancy"}',
                                  '{"problems": []}', '[1, 2]', '{"seed": -3}', '{"pp": 2}'])
def test_errors(text, tmp_path):
    p = tmp_path / "bad.json"
    p.write_text(text)
    n = ctypes.c_int(0)
    assert mahyco_amd.library().mhy_read_json(str(p).encode(), None, 0, ctypes.byref(n)) == abi.MHY_ERR_IO


def test_capacity(tmp_path):
    p = tmp_path / "e.json"
    p.write_text(json.dumps({"problems": [{"nx": 8}, {"nx": 9}, {"nx": 10}]}))
    L = mahyco_amd.library()
    n = ctypes.c_int(0)
    assert L.mhy_read_json(str(p).encode(), None, 0, ctypes.byref(n)) == abi.MHY_OK and n.value == 3
    arr = (abi.MhyConfig * 2)()
    assert L.mhy_read_json(str(p).encode(), arr, 2, ctypes.byref(n)) == abi.MHY_ERR_ARG
    assert L.mhy_read_json(str(tmp_path / "missing.json").encode(), None, 0, ctypes.byref(n)) == abi.MHY_ERR_IO
