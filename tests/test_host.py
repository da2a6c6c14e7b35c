"""Host-side checks that run without a GPU: the C-ABI library loads and exports
every entry point include/phx.h declares, configuration validation mirrors the
reference's rejections, the CLI keeps the reference's process contract, and the
data loaders read the reference formats."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "phx.h")).read()
    return sorted(set(re.findall(r"\b(phx_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(gpu_lib):
    names = _declared()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(gpu_lib, n)]
    assert not missing, missing


def test_package_lists_exports():
    import parahpop_amd
    assert sorted(parahpop_amd.EXPORTED) == _declared()


def _write_cfg(tmp, model, **ex):
    from parahpop_amd import synth
    with open(os.path.join(tmp, "constants.json"), "w") as f:
        json.dump(synth.CONSTANTS, f)
    cfg = {"run": {"duration_seconds": 600.0},
           "execution": {"maxThreads": 1, "integration": {"initialStepSize": 60.0, "maxNumSteps": 100}},
           "model": model}
    cfg["execution"].update(ex)
    p = os.path.join(tmp, "CONFIG.json")
    with open(p, "w") as f:
        json.dump(cfg, f)
    return p


def _create_error(path):
    import ctypes
    import parahpop_amd
    lib = parahpop_amd.library()
    h = ctypes.c_void_p()
    rc = lib.phx_create(path.encode(), ctypes.byref(h))
    if rc == 0:
        lib.phx_destroy(h)
        return rc, ""
    return rc, lib.phx_last_error(None).decode()


@pytest.mark.parametrize("model,expect", [
    ({"environment": {"bodies": {"Earth": {"gravity": True}}, "constants": "constants.json", "bogus": 1}}, "Unsupported model.environment option"),
    ({"environment": {"bodies": {"Sun": {"gravity": True}}, "constants": "constants.json"}}, "Earth must be configured"),
    ({"environment": {"bodies": {"Earth": {"gravity": {"shape": {"file": "missing.tab"}}}}, "constants": "constants.json"}}, "Required data file not found"),
    ({"environment": {"bodies": {"Earth": {"gravity": True}}, "ephemeris": ["de440s.brie"], "constants": "constants.json"}},
     "Required data file not found"),
    ({"environment": {"bodies": {"Earth": {"gravity": True}}, "orientations": ["itrf.brot2"], "constants": "constants.json"}},
     "Required data file not found"),
    ({"environment": {"bodies": {"Earth": {"gravity": True}, "Moon": {"radiation": True}}, "constants": "constants.json"}}, "Radiation is supported for the Sun only"),
    ({"environment": {"bodies": {"Earth": {"gravity": True}, "Vulcan": {"gravity": True}}, "constants": "constants.json"}}, "Unsupported body"),
    ({"environment": {"bodies": {"Earth": {"gravity": True}}, "constants": "constants.json",
                      "earthCorrections": {"model": "iers2010_fes2004", "solidEarthTides": True}}}, "eopFile"),
    ({"environment": {"bodies": {"Earth": {"atmosphere": {"exponential": [
        {"rho0": 1.0, "h0": 100.0, "scaleHeight": 10.0, "hCutoff": 300.0},
        {"rho0": 1.0, "h0": 200.0, "scaleHeight": 10.0, "hCutoff": 400.0}]}}}, "constants": "constants.json"}},
     "Overlapping atmosphere segments"),
    ({"environment": {"bodies": {"Earth": {"atmosphere": {"exponential": {"rho0": 1.0, "h0": 100.0, "scaleHeight": 10.0}}}},
                      "constants": "constants.json"}}, "hCutoff not specified"),
    ({"environment": {"bodies": {"Earth": {"atmosphere": {"exponential": {"rho0": 1.0, "h0": 100.0, "scaleHeight": 10.0,
                                                                          "hCutoff": 200.0, "T0": 1.0}}}},
                      "constants": "constants.json"}}, "Invalid atmosphere field"),
    ({"environment": {"bodies": {"Earth": {"atmosphere": {"exponential": {"rho0": -1.0, "h0": 100.0, "scaleHeight": 10.0,
                                                                          "hCutoff": 200.0}}}},
                      "constants": "constants.json"}}, "rho0 must be non-negative"),
    ({"environment": {"bodies": {"Earth": {"atmosphere": {"exponential": [], "nrlmsise00": {}}}},
                      "constants": "constants.json"}}, "single-key dispatch"),
    ({"environment": {"bodies": {"Earth": {"atmosphere": {"jacchia": {}}}}, "constants": "constants.json"}},
     "Unknown atmosphere model"),
    ({"environment": {"bodies": {"Earth": {"gravity": True}, "Moon": {"gravity": {"oblateness": True}}},
                      "constants": "constants.json"}}, "needs its rotation data"),
    ({"environment": {"bodies": {"Earth": {"gravity": True}, "Moon": {"gravity": {"oblateness": True}, "radius": 0.0}},
                      "constants": "constants.json"}}, "Body field needs the body radius"),
    ({"environment": {"bodies": {"Earth": {"gravity": ["oblateness", "rings"]}}, "constants": "constants.json"}},
     "Unsupported gravity feature"),
    ({"environment": {"bodies": {"Earth": {"gravity": True, "flattening": 1.0}}, "constants": "constants.json"}},
     "flattening must lie in [0, 1)"),
    ({"environment": {"bodies": {"Earth": {"atmosphere": {"exponential": [
        {"rho0": 1.0, "h0": 10.0 * k, "scaleHeight": 10.0, "hCutoff": 10.0 * k + 10.0} for k in range(33)]}}},
        "constants": "constants.json"}}, "at most 32 exponential atmosphere bands"),
])
def test_config_rejections(tmp_path, model, expect):
    rc, msg = _create_error(_write_cfg(str(tmp_path), model))
    assert rc == -1 and expect in msg, msg


def test_execution_rejections(tmp_path):
    model = {"environment": {"bodies": {"Earth": {"gravity": True}}, "constants": "constants.json"}}
    p = _write_cfg(str(tmp_path), model)
    cfg = json.load(open(p))
    for bad, expect in [({"maxNumSteps": 0}, "maxNumSteps and maxThreads"), ({"initialStepSize": -1.0}, "initialStepSize"),
                        ({"absTol": 1e-9}, "Unsupported integration option")]:
        c = json.loads(json.dumps(cfg))
        c["execution"]["integration"].update(bad)
        json.dump(c, open(p, "w"))
        rc, msg = _create_error(p)
        assert rc == -1 and expect in msg, msg
    c = json.loads(json.dumps(cfg))
    c["execution"]["maxThreads"] = -1
    json.dump(c, open(p, "w"))
    rc, msg = _create_error(p)
    assert rc == -1 and "maxNumSteps and maxThreads" in msg
    for prec, expect in [({"mode": "half"}, "precision.mode"), ({"mode": "mixed", "fp64Degree": 0}, "fp64Degree must be"),
                         ({"mode": "mixed", "fp64Order": -1}, "fp64Order must be"),
                         ({"mode": "mixed", "fp32Degree": 30}, "fp32Degree")]:
        c = json.loads(json.dumps(cfg))
        c["execution"]["precision"] = prec
        json.dump(c, open(p, "w"))
        rc, msg = _create_error(p)
        assert rc == -1 and expect in msg, msg


def test_cli_contract(tmp_path):
    exe = os.path.join(ROOT, "parahpop_amd", "phx_gpu")
    out = subprocess.run([exe, "--help"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert out.returncode == 0 and "CONFIG.json TARGETS.json OUTPUT_DIRECTORY" in out.stdout
    out = subprocess.run([exe, "a"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert out.returncode == 1 and out.stderr.startswith("ParaHPOP: Expected CONFIG.json")
    d = tmp_path / "out"
    d.mkdir()
    (d / "summary.json").write_text("{}")
    out = subprocess.run([exe, "c.json", "t.json", str(d)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert out.returncode == 1 and "Output exists" in out.stderr


def test_loaders_read_reference_formats(workdir):
    """The oracle context parses the synthetic gravity, SW-All and EOP files the
    generator writes in the reference formats (shared host loader)."""
    import oracle
    from parahpop_amd import synth
    cfg, _ = synth.make_workload(str(workdir / "loaders"), 2, n=4, duration=600.0)
    o = oracle.Oracle(cfg)
    assert o.h
    tab = open(os.path.join(os.path.dirname(cfg), "gravity_70.tab")).read().splitlines()
    assert len(tab) == 1 + sum(n + 1 for n in range(1, 71))
    assert float(tab[3].split(",")[2]) == -4.8417e-4  # (2, 0) row


def test_population_distributions():
    from parahpop_amd import synth
    pop = synth.population(20000, 2, mix=(0.8, 0.2, 0.0))
    r = np.linalg.norm(pop["state"][:3], axis=0)
    leo = pop["kind"] == 0
    assert abs(leo.mean() - 0.8) < 0.02
    # perigee >= 150 km; apogee <= a (1 + e) with a <= R_E + 2000 km and e <= 0.02
    assert np.all(r[leo] > synth.R_EARTH + 150) and np.all(r[leo] <= (synth.R_EARTH + 2000.0) * 1.02)
    am = pop["drag_area"] / pop["mass"]
    assert abs(np.median(np.log(am)) - np.log(0.01)) < 0.05


def test_json_writer_formats_numbers_as_nlohmann(tmp_path):
    """summary.json and configuration_resolved.json are written by nlohmann::json
    in the reference (driver.cpp:166-175): shortest round-trip digits, decimal
    notation for decimal exponents in (-4, 15], else d.ddde+XX; keys sorted.
    Expected texts are nlohmann::json::dump() outputs of the same values."""
    cases = [(600.0, "600.0"), (0.1, "0.1"), (1e-5, "1e-05"), (2.5e-5, "2.5e-05"), (1.5e20, "1.5e+20"),
             (1e15, "1e+15"), (1e14, "100000000000000.0"), (-0.0, "-0.0"), (0.0001, "0.0001"),
             (0.001, "0.001"), (3.141592653589793, "3.141592653589793"), (86400.0, "86400.0"),
             (4035.0416666666665, "4035.0416666666665"), (5e-324, "5e-324"),
             (1.7976931348623157e308, "1.7976931348623157e+308"), (-123.456, "-123.456"), (1e100, "1e+100")]
    src = tmp_path / "fmt.cpp"
    body = "\n".join('    std::printf("%%s\\n", phx::json::numberText(phx::json::Value::of(%r)).c_str());' % v
                     for v, _ in cases)
    src.write_text('#include "json_lite.h"\n#include <cstdio>\nint main() {\n%s\n'
                   '    phx::json::Value o = phx::json::parse("{\\"b\\": 1, \\"a\\": {\\"d\\": [2, 1.0], \\"c\\": -3}}");\n'
                   '    std::printf("%%s\\n", phx::json::dump(phx::json::sortedKeys(o)).c_str());\n}\n' % body)
    exe = tmp_path / "fmt"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "parahpop_amd", "csrc"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, stdout=subprocess.PIPE, text=True).stdout.split("\n")
    assert out[:len(cases)] == [t for _, t in cases]
    tail = "\n".join(out[len(cases):]).strip()
    assert tail == '{\n  "a": {\n    "c": -3,\n    "d": [\n      2,\n      1.0\n    ]\n  },\n  "b": 1\n}'
