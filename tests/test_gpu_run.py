"""The reference program's runs from configuration files on the GPU
(src/main.cpp:163-1100 with src/IO/ProgramOptions.cpp)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


@pytest.fixture(scope="module")
def V():
    import inovesa_amd as V
    return V


def _rows(path, dset, tmp):
    out = os.path.join(tmp, "d.bin")
    subprocess.run([H5DUMP, "-d", dset, "-b", "LE", "-o", out, path],
                   capture_output=True, check=True, timeout=60)
    return np.fromfile(out, np.float32)


def _opts(tmp_path, extra):
    from inovesa_amd.options import ProgramOptions
    cfg = tmp_path / "run.cfg"
    cfg.write_text("GridSize=128\nStepsPerTs=200\nrotations=0.1\n"
                   "VacuumGap=0.032\nLinearRF=false\noutstep=5\n" + extra)
    return ProgramOptions().parse(["-c", str(cfg)])


def test_single_bunch_config_runs_fused_with_results(V, tmp_path):
    from inovesa_amd.run import FusedRun, main, simulation
    out = str(tmp_path / "out.h5")
    o = _opts(tmp_path, f"SavePhaseSpace=2\noutput={out}\n")
    sim = simulation(o)
    # laststep = ceil(steps * rotations) with rotations held as float
    # (src/main.cpp:296, 355): 200 * 0.1f = 20.0000003 -> 21
    assert isinstance(sim, FusedRun) and sim.laststep == 21
    st = sim.run()
    # outputs at steps 0, 5, 10, 15, 20 and the final one
    t = _rows(out, "/Info/AxisValues_t", str(tmp_path))
    np.testing.assert_allclose(t, np.array([0, 5, 10, 15, 20, 21]) / 200.0)
    # the phase space at outputs 0, 2, 4 and the final one
    assert len(_rows(out, "/PhaseSpace/axis0", str(tmp_path))) == 4
    attrs = subprocess.run([H5DUMP, "-a", "/Info/Parameters/GridSize", out],
                           capture_output=True, text=True).stdout
    assert "H5T_STD_U32LE" in attrs and "128" in attrs
    # the same run through the command line entry point
    out2 = str(tmp_path / "cli.h5")
    assert main(["-c", str(tmp_path / "run.cfg"), "-o", out2,
                 "--SavePhaseSpace", "2"]) == 0
    assert os.path.exists(out2 + ".cfg")
    from inovesa_amd.hdf5 import readPhaseSpace
    a, _ = readPhaseSpace(out, -1)
    b, _ = readPhaseSpace(out2, -1)
    assert np.array_equal(a, b)
    assert np.array_equal(a[0], st.mesh()[0])


def test_map_loop_equals_per_map_step(V, tmp_path):
    """The loop of src/main.cpp:948-1053 driven map by map equals the step
    object's per-map path bit for bit."""
    from inovesa_amd.run import MapLoop, Setup
    o = _opts(tmp_path, "")
    s = Setup(o)
    loop = MapLoop(s)
    init = loop.t1.getData()[0]
    st = V.Step(dict(s.step_desc(), fused=0, renormalize=0), init, s.zwake,
                s.zrad)
    loop.run(12)
    st.run(12)
    mesh, proj, _ = st.mesh()
    assert np.array_equal(loop.t1.getData()[0], mesh)
    assert np.array_equal(loop.t1.getProjection(0)[0], proj)


def test_multibunch_config(V, tmp_path):
    """Three buckets, two filled: both bunches in one padded transform of
    upper_power_of_two(ceil(n * 3 * spacing)) bins, written per bunch."""
    from inovesa_amd.run import MapLoop, simulation
    out = str(tmp_path / "mb.h5")
    o = _opts(tmp_path, f"GridSize=32\nBunchCurrent=1e-3\nBunchCurrent=0\n"
                        f"BunchCurrent=5e-4\noutput={out}\n")
    sim = simulation(o)
    assert isinstance(sim, MapLoop) and sim.s.nb == 2
    assert sim.s.wake_bins == 4096
    sim.run(10)
    wp = _rows(out, "/WakePotential/data", str(tmp_path)).reshape(-1, 2, 32)
    assert wp.shape[0] == 3 and np.all(np.isfinite(wp))
    pop = _rows(out, "/BunchPopulation/data", str(tmp_path)).reshape(-1, 2)
    np.testing.assert_allclose(pop[0], [2 / 3, 1 / 3], rtol=1e-4)
    bk = subprocess.run([H5DUMP, "-d", "/Info/BucketNumbers", out],
                        capture_output=True, text=True).stdout
    assert "(0): 2, 0" in bk


def test_dynamic_rf_config(V, tmp_path):
    """RF phase and amplitude noise make the RF kick a DynamicRFKickMap;
    every applied modulation goes to /RFKicks/data."""
    from inovesa_amd.run import MapLoop, simulation
    out = str(tmp_path / "rf.h5")
    o = _opts(tmp_path, f"GridSize=64\nRFPhaseSpread=0.5\n"
                        f"RFAmplitudeSpread=1e-3\noutput={out}\n")
    sim = simulation(o)
    assert isinstance(sim, MapLoop) and sim.drfm is not None
    sim.run(10)
    kicks = _rows(out, "/RFKicks/data", str(tmp_path)).reshape(-1, 2)
    assert kicks.shape[0] == 10
    assert np.std(kicks[:, 1]) > 0
