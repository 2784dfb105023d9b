"""The reference's JSON configuration, initial conditions and VTI output as
restated in include/smac_config.hpp (host only, no GPU), checked through
tests/cpp/config_tool against independent Python readings (tests/config_ref.py):
every parsed value, the initial alpha and v fields, the VTI header and blocks,
and the loader's error texts for the reference's failure cases
(CFD/config/ConfigLoader.cpp, CFD/config/SimulationConfig.cpp)."""
import copy
import json
import os
import subprocess

import numpy as np
import pytest

import config_ref
from cpp import drivers as cpp_drivers

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
REF_CONFIG = "/root/reference/CFD/config.json"


@pytest.fixture(scope="module")
def tool():
    return cpp_drivers.config_tool()


def _run(tool, cfg, tmp_path):
    path = os.path.join(tmp_path, "config.json")
    with open(path, "w") as f:
        json.dump(cfg, f)
    return subprocess.run([tool, path, str(tmp_path)], capture_output=True, text=True, timeout=60)


def _load():
    with open(os.path.join(GOLDEN, "dam_break_config.json")) as f:
        return json.load(f)


def test_values_fields_and_vti(tool, tmp_path):
    cfg = _load()
    cfg["grid"]["cells"] = [12, 8, 6]
    cfg["grid"]["origin"] = [0.25, -0.25, 0.0]
    cfg["boundary_conditions"].append({"id": 3, "type": "inflow", "velocity": [0.5, 0.0, -0.25]})
    cfg["boundary_faces"][1]["boundary_id"] = 3
    r = _run(tool, cfg, tmp_path)
    assert r.returncode == 0, r.stderr
    kv = dict(line.split("=", 1) for line in r.stdout.strip().splitlines())
    num = lambda s: [float(t) for t in s.split(",")]
    assert kv["cells"] == "12,8,6"
    assert num(kv["size"]) == cfg["grid"]["size"] and num(kv["origin"]) == cfg["grid"]["origin"]
    assert num(kv["rho"]) == [1.0, 1000.0] and num(kv["mu"]) == [1.8e-5, 1.0e-3]
    assert num(kv["gravity"]) == cfg["fluid"]["gravity"]
    t = cfg["time"]
    assert num(kv["time"]) == [t["initial_dt"], t["maximum_dt"], t["end_time"], t["interval"], t["cfl"], t["vof_cfl"]]
    assert kv["gmg_levels"] == "4" and float(kv["tol"]) == 1e-12
    assert kv["num_boundary_ids"] == "4"            # max id 3, plus row 0
    assert kv["pure_neumann"] == "0"                # an inflow row makes it non pure-Neumann
    assert kv["alpha_substeps"] == "2"              # ceil(0.2 / 0.1)
    d = config_ref.desc(cfg)
    assert [int(s) for s in kv["side_bcid"].split(",")] == list(d["side_bcid"])
    for row, ent in enumerate(d["bc_table"]):
        vals = num(kv[f"bc{row}"])
        assert int(vals[0]) == ent[0] and vals[1:] == list(ent[1:])
    alpha, v = config_ref.initial_fields(cfg)
    ta = np.fromfile(os.path.join(tmp_path, "alpha.f64")).reshape(alpha.shape)
    tv = np.fromfile(os.path.join(tmp_path, "v.f64")).reshape(v.shape)
    assert alpha.sum() > 0 and (v != 0).any()
    np.testing.assert_array_equal(ta, alpha)
    np.testing.assert_array_equal(tv, v)
    attrs, arrs = config_ref.read_vti(os.path.join(tmp_path, "result_000007.vti"))
    nx, ny, nz = 12, 8, 6
    assert attrs["extent"] == [0, nx, 0, ny, 0, nz] and attrs["tail_ok"]
    assert attrs["origin"] == [0.25, -0.25, 0.0]
    assert attrs["spacing"] == pytest.approx([2.0 / nx, 1.0 / ny, 1.0 / nz], rel=1e-15)
    assert attrs["origin_text"].split()[0] == "%.15e" % 0.25
    k, j, i = np.meshgrid(np.arange(1, nz + 1), np.arange(1, ny + 1), np.arange(1, nx + 1), indexing="ij")
    np.testing.assert_array_equal(arrs["pressure"][0], (i + 1000.0 * j + 1e6 * k).astype(np.float32).ravel())
    np.testing.assert_array_equal(arrs["alpha"][0], alpha[1:-1, 1:-1, 1:-1].astype(np.float32).ravel())
    vel, comps = arrs["velocity"]
    assert comps == 3
    vel = vel.reshape(nz, ny, nx, 3)
    vc = (0.5 * (v[1:-1, 1:-2, 1:-1] + v[1:-1, 2:-1, 1:-1])).astype(np.float32)
    np.testing.assert_array_equal(vel[..., 1], vc)
    assert not vel[..., 0].any() and not vel[..., 2].any()
    assert not arrs["ibm_solid_fraction"][0].any()


def _err(tool, cfg, tmp_path):
    r = _run(tool, cfg, tmp_path)
    assert r.returncode == 2, (r.stdout, r.stderr)
    return r.stderr.strip()


def test_loader_errors(tool, tmp_path):
    base = _load()
    cases = []
    c = copy.deepcopy(base)
    del c["grid"]["origin"]
    cases.append((c, 'configuration error at "grid.origin": required value is missing'))
    c = copy.deepcopy(base)
    c["grid"]["cells"] = [32, 16.0, 16]
    cases.append((c, 'configuration error at "grid.cells[1]": must be an integer'))
    c = copy.deepcopy(base)
    c["grid"]["cells"] = [32, 16]
    cases.append((c, 'configuration error at "grid.cells": must contain exactly three integers'))
    c = copy.deepcopy(base)
    c["boundary_conditions"].append({"id": 1, "type": "slip"})
    cases.append((c, "configuration error: boundary condition ID 1 is defined more than once"))
    c = copy.deepcopy(base)
    c["boundary_faces"][0]["boundary_id"] = 5
    cases.append((c, "configuration error: boundary face refers to undefined boundary ID 5"))
    c = copy.deepcopy(base)
    c["boundary_faces"][1]["side"] = "min"
    cases.append((c, "configuration error: the same outer boundary face is assigned more than once"))
    c = copy.deepcopy(base)
    c["boundary_conditions"][0]["type"] = "wall"
    cases.append((c, 'configuration error at "boundary_conditions[0].type": unknown boundary type "wall"'))
    c = copy.deepcopy(base)
    c["time"]["initial_dt"] = 1.0
    cases.append((c, "configuration error: time.initial_dt must not exceed time.maximum_dt"))
    c = copy.deepcopy(base)
    c["initial_conditions"][0]["alpha"] = 1.5
    cases.append((c, "configuration error: VOF alpha must be between zero and one"))
    c = copy.deepcopy(base)
    c["initial_conditions"][1]["type"] = "vof_cube"
    cases.append((c, 'configuration error at "initial_conditions[1].type": unknown initial condition type "vof_cube"'))
    c = copy.deepcopy(base)
    c["hardware"]["use_gpu"] = 1
    cases.append((c, 'configuration error at "hardware.use_gpu": must be true or false'))
    for cfg, msg in cases:
        assert _err(tool, cfg, tmp_path) == msg


def test_malformed_json(tool, tmp_path):
    path = os.path.join(tmp_path, "bad.json")
    with open(path, "w") as f:
        f.write('{"hardware": {"use_gpu": true,}}')
    r = subprocess.run([tool, path, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and r.stderr.startswith("JSON parse error at line 1"), r.stderr


def test_json_reader_escapes_and_numbers(tool, tmp_path):
    cfg = _load()
    cfg["output"]["directory"] = "r\u00e9sultats/\"q\"\\x\ty"
    cfg["time"]["end_time"] = 2.5e-2
    r = _run(tool, cfg, tmp_path)
    assert r.returncode == 0, r.stderr
    kv = dict(line.split("=", 1) for line in r.stdout.strip().splitlines())
    assert kv["directory"] == cfg["output"]["directory"]
    assert float(kv["time"].split(",")[2]) == 2.5e-2


@pytest.mark.skipif(not os.path.exists(REF_CONFIG), reason="the reference tree is not present")
def test_reference_config_fails_as_the_reference_loader(tool, tmp_path):
    """The reference's own CFD/config.json has no grid.origin, which its loader
    requires (ConfigLoader.cpp:375-383): the same error as the reference."""
    r = subprocess.run([tool, REF_CONFIG, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert r.stderr.strip() == 'configuration error at "grid.origin": required value is missing'
    with open(REF_CONFIG) as f:
        cfg = json.load(f)
    cfg["grid"]["origin"] = [0.0, 0.0, 0.0]
    r = _run(tool, cfg, tmp_path)
    assert r.returncode == 0, r.stderr
    alpha, v = config_ref.initial_fields(cfg)
    ta = np.fromfile(os.path.join(tmp_path, "alpha.f64")).reshape(alpha.shape)
    np.testing.assert_array_equal(ta, alpha)
