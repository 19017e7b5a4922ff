"""The C++ interface (include/inovesa_amd.hpp) through the C++ test program
tests/cpp/test_reference_api.cpp: the reference's unit-test vectors and the
reference time loop written with the reference-shaped classes."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "cpp"))
import build as cpp_build  # noqa: E402


def _arr(key, values):
    vals = [float(v) for v in np.ravel(values)]
    return f"{key} {len(vals)} " + " ".join(repr(v) for v in vals)


def write_fixture(path, golden):
    g = golden("reference_vectors.json")
    lines = []
    for name in ("kickmap_x", "kickmap_x_overflow", "kickmap_y"):
        r = g[name]
        lines.append(" ".join([name, _arr("axis", [r["axis"]]),
                               _arr("it", [r["it"]]),
                               _arr("offset", r["offset"]),
                               _arr("data_in", r["data_in"]),
                               _arr("data_out", r["data_out"])]))
    for name in ("rotation_linear_noclamp_map", "rotation_linear_offcenter",
                 "rotation_quadratic_noclamp_nomap",
                 "rotation_cubic_clamp_map"):
        r = g[name]
        lines.append(" ".join([name] + [_arr(k, [float(r[k])]) for k in
                                        ("qmin", "qmax", "pmin", "pmax",
                                         "angle", "it", "clamp",
                                         "rotmapsize")]
                              + [_arr("data_in", r["data_in"]),
                                 _arr("data_out", r["data_out"])]))
    for i, r in enumerate(g["rotation_throws"]):
        lines.append(" ".join([f"rotation_throws_{i}", _arr("it", [r["it"]]),
                               _arr("clamp", [float(r["clamp"])]),
                               _arr("rotmapsize", [r["rotmapsize"]])]))
    r = g["phasespace_gaussian"]
    lines.append(" ".join(["phasespace_gaussian"] + [
        _arr(k, [r[k]]) for k in ("n", "min", "max", "integral", "sigma",
                                  "rel_tol")]))
    r = g["impedance_addition"]
    lines.append(" ".join(["impedance_addition", _arr("z1", r["z1"]),
                           _arr("z2", r["z2"])]))
    r = g["const_impedance"]
    lines.append(" ".join(["const_impedance", _arr("n", [r["n"]]),
                           _arr("z", [r["z"], 0.0]),
                           _arr("expect", [[v, 0.0] for v in r["expect"]])]))
    # the time loop: a 64x64 kick-drift configuration
    from inovesa_amd.workloads import initial_mesh, kick_drift_config
    desc, d = kick_drift_config(n=64, steps_per_ts=500)
    keys = ("n", "qmin", "qmax", "pmin", "pmax", "qscale", "pscale",
            "revolutionpart", "V_RF", "f_RF", "V0", "E0", "e1",
            "spacing_bins", "wakescaling")
    rec = ["time_loop", _arr("steps", [12])]
    rec += [_arr(k, [desc[k]]) for k in keys]
    rec += [_arr("slip", desc["slip"]), _arr("nmax", [d["padded_bins"]]),
            _arr("f0", [np.float32(d["f0"])]), _arr("fmax", [d["fmax"]]),
            _arr("gap", [d["ring"]["vacuum_gap"]]),
            _arr("init", initial_mesh(64, seed=3))]
    lines.append(" ".join(rec))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def test_interface_compiles():
    """The header and the test program compile (no device needed)."""
    r = subprocess.run(cpp_build.command(syntax_only=True), capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_reference_unit_tests_and_time_loop(tmp_path, golden):
    if not os.path.exists(cpp_build.BIN):
        pytest.fail(f"{cpp_build.BIN} is missing: build with "
                    "`python tests/cpp/build.py` (after the library)")
    fix = tmp_path / "fixture.txt"
    write_fixture(str(fix), golden)
    print("running", cpp_build.BIN, flush=True)
    r = subprocess.run([cpp_build.BIN, str(fix)], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
    for case in ("kickmap_x ok", "rotation_cubic_clamp_map ok",
                 "time_loop ok", "electric_field ok", "impedance ok"):
        assert case in r.stdout
