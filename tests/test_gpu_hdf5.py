"""The step's output files and resume (src/main.cpp:908-997, 1079-1099;
src/IO/HDF5File.cpp), the outstep moments, and RenormalizeCharge
(src/main.cpp:953-955) on the GPU."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from _helpers import oracle_cfg, rel_linf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def V():
    import inovesa_amd as V
    return V


def _setup(n=256, **kw):
    from inovesa_amd.workloads import impedances, initial_mesh, kick_drift_config
    desc, d = kick_drift_config(n=n, steps_per_ts=500, **kw)
    zw, zc = impedances(d)
    return desc, d, zw, zc, initial_mesh(n, seed=9)


def _h5read(path, dset, tmp):
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    out = os.path.join(tmp, "d.bin")
    subprocess.run([h5dump, "-d", dset, "-b", "LE", "-o", out, path],
                   capture_output=True, check=True, timeout=60)
    return np.fromfile(out, np.float32)


def test_moments_match_oracle(V, orc):
    """integrate(), variance(0), updateYProjection(), variance(1) of the
    step's state against the oracle's float restatement of
    src/PS/PhaseSpace.cpp:241-340 (the integral here is FP64)."""
    desc, d, zw, zc, init = _setup()
    st = V.Step(desc, init, zw, zc)
    st.run(7)
    m = st.moments()
    mesh, proj, _ = st.mesh()
    n = desc["n"]
    ax0, d0, _ = orc.ruler(n, desc["qmin"], desc["qmax"])
    ax1, d1, _ = orc.ruler(n, desc["pmin"], desc["pmax"])
    ws = orc.simpson_weights(n, d0)
    data = mesh.reshape(1, n, n)
    xp = orc.project_x(data, ws)
    yp = orc.project_y(data, ws)
    fill, _ = orc.integrate(xp, ws)
    assert np.array_equal(m["xproj"], xp[0])
    assert rel_linf(m["yproj"], yp[0]) <= 1e-6
    assert abs(m["population"] - fill[0]) <= 1e-6 * abs(fill[0])
    mx, vx, _ = orc.variance(xp, ax0, d0, [1.0], fill)
    my, vy, _ = orc.variance(yp, ax1, d1, [1.0], fill)
    assert abs(m["position"] - mx[0]) <= 1e-5 * max(1.0, abs(mx[0]))
    assert abs(m["length"] - np.sqrt(vx[0])) <= 1e-5 * np.sqrt(vx[0])
    assert abs(m["average"] - my[0]) <= 1e-5 * max(1.0, abs(my[0]))
    assert abs(m["spread"] - np.sqrt(vy[0])) <= 1e-5 * np.sqrt(vy[0])


def test_writer_outputs_and_resume(V, tmp_path):
    """StepWriter at every 10th step with the phase space at every 2nd
    output: the rows the reference loop writes, profiles equal to the step's
    own history, CSR power equal to the step's within 1e-5, and a resume
    from the last saved phase space that continues bit for bit."""
    from inovesa_amd.hdf5 import HDF5File, StepWriter, readPhaseSpace
    desc, d, zw, zc, init = _setup()
    st = V.Step(desc, init, zw, zc)
    st.record(64, profiles=True)
    path = str(tmp_path / "run.h5")
    w = StepWriter(path, st, desc, d, zw, zc, outstep=10, h5save=2)
    w.run(40)
    w.finish()
    csr, prof = st.history()
    rows = lambda ds: len(_h5read(path, ds, str(tmp_path)))
    n = desc["n"]
    assert rows("/Info/AxisValues_t") == 5     # steps 0, 10, 20, 30 + final
    assert rows("/PhaseSpace/axis0") == 3      # outputs 0 and 2 + final
    t = _h5read(path, "/Info/AxisValues_t", str(tmp_path))
    np.testing.assert_allclose(t, np.array([0, 10, 20, 30, 40]) / 500.0)
    bp = _h5read(path, "/BunchProfile/data", str(tmp_path)).reshape(5, n)
    for k, s in enumerate((10, 20, 30, 40)):
        # the profile written at step s is the one after s steps
        assert np.array_equal(bp[k + 1], prof[s - 1]), s
    ci = _h5read(path, "/CSR/Intensity/data", str(tmp_path))
    for k, s in enumerate((10, 20, 30)):
        # the step's CSR power of step s comes from the same profile
        assert abs(ci[k + 1] - csr[s]) <= 1e-5 * abs(csr[s]), (s, ci, csr)
    sp = _h5read(path, "/CSR/Spectrum/data", str(tmp_path))
    assert sp.size == 5 * zc.n // 2 and np.all(np.isfinite(sp))
    wp = _h5read(path, "/WakePotential/data", str(tmp_path)).reshape(5, n)
    assert np.abs(wp).max() > 0
    mesh_end = st.mesh()[0]
    ps, nsaved = readPhaseSpace(path, -1)
    assert nsaved == 3 and np.array_equal(ps[0], mesh_end)
    # resume (src/PS/PhaseSpaceFactory.cpp:19-53) and continue
    resumed = V.Step(desc, ps[0], zw, zc)
    st.run(25)
    resumed.run(25)
    for x, y in zip(st.mesh(), resumed.mesh()):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("every", [1, 5])
def test_renormalize_charge_vs_oracle(V, orc, every):
    """RenormalizeCharge: before the maps of every step whose count is a
    multiple of `every`, the mesh is scaled to unit charge (the profile
    that forms the wake stays).  The factor divides by the FP64 integral
    here and by the reference's sequential float sum of n terms there,
    whose own rounding reaches about 1e-6: per step, from the same state,
    within the 1e-5 bar (and the profile within 1e-5), and a charge of 1
    after each due step."""
    desc, d, zw, zc, init = _setup()
    desc = dict(desc, renormalize=every)
    init = (init * np.float32(1.05)).astype(np.float32)   # charge 1.05
    st = V.Step(desc, init, zw, zc)
    st.record(16, profiles=True)
    sim = orc.sim(oracle_cfg(desc, zw.n), init, zw.data(), zc.data())
    free = orc.sim(oracle_cfg(desc, zw.n), init, zw.data(), zc.data())
    n = desc["n"]
    for k in range(12):
        st.run(1)
        sim.run(1)
        free.run(1)
        mesh, proj, wake = st.mesh()
        omesh, oproj, owake = sim.get()
        # per step, from the same state
        assert rel_linf(mesh, omesh) <= 1e-5, k
        assert rel_linf(proj, oproj) <= 1e-5, k
        sim.set_mesh(mesh.reshape(n, n))
    # free runs over 12 steps
    assert rel_linf(st.mesh()[0], free.get()[0]) <= 1e-4
    # a step that renormalised starts from unit charge: the next profile
    # integrates to 1 up to the maps' own change of charge
    fill = st.moments()["population"]
    plain = V.Step(dict(desc, renormalize=0), init, zw, zc)
    plain.run(12)
    assert abs(fill - 1.0) < abs(plain.moments()["population"] - 1.0)
    assert abs(fill - 1.0) < 1e-3


@pytest.mark.parametrize("strip", [0, 8])
def test_renormalize_batch_equals_single(V, strip):
    """The batched form (tiles, or strips of 8 tiles) renormalises exactly
    as the single step."""
    desc, d, zw, zc, init = _setup(n=128)
    desc = dict(desc, renormalize=3)
    b = V.Batch([desc, desc], [init, init * np.float32(0.9)], [zw, zw],
                [zc, zc])
    if strip:
        b.retile_size(128, 2, 256, strip=strip)
    b.run(10)
    for k, x0 in enumerate((init, init * np.float32(0.9))):
        s = V.Step(desc, x0, zw, zc)
        s.run(10)
        for x, y in zip(b.mesh(k), s.mesh()):
            assert np.array_equal(x, y), k


def test_renormalize_strip_with_staged_first_window(V):
    """Strips of 8 tiles of 4 rows at 512^2: the rows of each strip's first
    drift window are staged before the renormalisation factor is known
    (their halo fits the two staging tiles) and scaled in place afterwards;
    every step still equals the single step's.  2000 steps per synchrotron
    period, as in the 512^2 benchmark workload, keep the drift's halo within
    a tile."""
    from inovesa_amd.workloads import impedances, initial_mesh, kick_drift_config
    desc, d = kick_drift_config(n=512, steps_per_ts=2000)
    zw, zc = impedances(d)
    init = initial_mesh(512, seed=9)
    desc = dict(desc, renormalize=2)
    inits = [init, init * np.float32(0.8)]
    b = V.Batch([desc, desc], inits, [zw, zw], [zc, zc])
    b.retile_size(512, 4, 512, strip=8)
    cfg = b.config(0)
    assert cfg["strip"] == 8 and cfg["tile_rows"] == 4
    assert cfg["stride"] - cfg["tile_rows"] <= cfg["tile_rows"]  # halo <= T
    b.run(9)
    for k, x0 in enumerate(inits):
        s = V.Step(desc, x0, zw, zc)
        s.run(9)
        for x, y in zip(b.mesh(k), s.mesh()):
            assert np.array_equal(x, y), k


def test_renormalize_needs_direct_fused_step(V):
    desc, d, zw, zc, init = _setup(n=128)
    with pytest.raises(V.VfpsInvalid):
        V.Step(dict(desc, renormalize=4, wake_method=1), init, zw, zc)
    with pytest.raises(V.VfpsInvalid):
        V.Step(dict(desc, renormalize=4, fused=0), init, zw, zc)
