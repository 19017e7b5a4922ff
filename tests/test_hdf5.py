"""HDF5 result files in the layout of src/IO/HDF5File.cpp (host only):
groups, data sets, shapes, unit attributes and axis links checked with the
HDF5 tools' h5dump; the phase space read back by readPhaseSpace."""
import os
import shutil
import subprocess

import numpy as np
import pytest

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


@pytest.fixture(scope="module")
def H():
    from inovesa_amd import hdf5
    try:
        hdf5.lib()
    except ImportError as e:
        pytest.skip(str(e))
    return hdf5


def _dump(path, *args):
    r = subprocess.run([H5DUMP, *args, path], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _read(path, dset, tmp):
    out = os.path.join(tmp, "d.bin")
    subprocess.run([H5DUMP, "-d", dset, "-b", "LE", "-o", out, path],
                   capture_output=True, check=True, timeout=60)
    return np.fromfile(out, np.float32)


def _file(H, path, nb=2, n=16, nmax=64, nparticles=3):
    ax = np.linspace(-6, 6, n, dtype=np.float32)
    faxis = np.arange(nmax // 2, dtype=np.float32) * np.float32(0.25)
    z = (np.arange(nmax) + 1j * np.arange(nmax) / 2).astype(np.complex64)
    return H.HDF5File(path, nb, n, ax, 2.0e-3, 6.11e5, 2.25e-5, 9e6, 3e-3,
                      3e-3 / 9e6, buckets=list(range(nb)), maxn=nmax // 2,
                      axis_f=faxis, hertz=1.5e11, volts=12.5,
                      factor4WattPerHertz=2e-12, factor4Watts=0.3,
                      impedance=z, nparticles=nparticles)


def test_layout_matches_reference(H, tmp_path):
    path = str(tmp_path / "out.h5")
    f = _file(H, path)
    f.addParameterToGroup("/Info", "ShieldingParameter", 0.893)
    f.addParameterToGroup("/Info/Parameters", "GridSize", 16)
    f.addParameterToGroup("/Info/Parameters", "ImpedanceFile", "none")
    f.close()
    head = _dump(path, "-H", "-p")
    # groups of HDF5File::_prepareFile (src/IO/HDF5File.cpp:558-587)
    for g in ("BunchLength", "BunchPopulation", "BunchPosition",
              "BunchProfile", "CSR", "Intensity", "Spectrum",
              "EnergyProfile", "EnergySpread", "EnergyAverage", "Particles",
              "PhaseSpace", "RFKicks", "Impedance", "Info", "Parameters",
              "SourceMap", "WakeFunction", "WakePotential"):
        assert f'GROUP "{g}"' in head, g
    # extendible data sets with their shapes (:44-149)
    for d, shape in (("/PhaseSpace/data", "( 0, 2, 16, 16 ) / ( H5S_UNLIMITED, 2, 16, 16 )"),
                     ("/BunchProfile/data", "( 0, 2, 16 ) / ( H5S_UNLIMITED, 2, 16 )"),
                     ("/CSR/Spectrum/data", "( 0, 2, 32 ) / ( H5S_UNLIMITED, 2, 32 )"),
                     ("/Particles/data", "( 0, 3, 2 ) / ( H5S_UNLIMITED, 3, 2 )"),
                     ("/RFKicks/data", "( 0, 2 ) / ( H5S_UNLIMITED, 2 )"),
                     ("/BunchProfile/padded", "( 0, 32 ) / ( H5S_UNLIMITED, 32 )"),
                     ("/Impedance/data/real", "( 32 ) / ( 32 )")):
        hd = _dump(path, "-H", "-d", d)
        assert shape in hd, (d, hd)
        if "H5S_UNLIMITED" in shape:
            assert "DEFLATE { LEVEL 6 }" in _dump(path, "-H", "-p", "-d", d)
    # soft links from each quantity's axes to /Info (:196-305)
    for link, target in (("/BunchProfile/axis0", "/Info/AxisValues_t"),
                         ("/BunchProfile/axis1", "/Info/AxisValues_z"),
                         ("/EnergyProfile/axis1", "/Info/AxisValues_E"),
                         ("/CSR/Spectrum/axis1", "/Info/AxisValues_f"),
                         ("/PhaseSpace/axis2", "/Info/AxisValues_E"),
                         ("/Impedance/axis0", "/Info/AxisValues_f")):
        assert f'SOFTLINK "{link.rsplit("/", 1)[1]}"' in head
        assert f'LINKTARGET "{target}"' in _dump(path, "-l", link)
    # unit attributes
    for d, attr in (("/Info/AxisValues_z", "Meter"),
                    ("/Info/AxisValues_t", "Turn"),
                    ("/BunchPopulation/data", "Coulomb"),
                    ("/WakePotential/data", "Volt"),
                    ("/CSR/Intensity/data", "Watt"),
                    ("/PhaseSpace/data", "AmperePerNBLPerNES")):
        assert f'ATTRIBUTE "{attr}"' in _dump(path, "-H", "-d", d), (d, attr)
    assert "(0): 2.25e-05" in _dump(path, "-a", "/Info/AxisValues_t/Second")
    assert "202.5" in _dump(path, "-a", "/Info/AxisValues_t/Turn")
    assert "0.893" in _dump(path, "-a", "/Info/ShieldingParameter")
    info = _dump(path, "-d", "/Info/BucketNumbers")
    assert "H5T_STD_U32LE" in info and "(0): 0, 1" in info
    assert "(0): 1, 1, 0" in _dump(path, "-d", "/Info/Inovesa_v")


def test_append_and_read_back(H, tmp_path):
    path = str(tmp_path / "run.h5")
    nb, n = 2, 16
    f = _file(H, path, nb=nb, n=n)
    rng = np.random.default_rng(4)
    meshes, profiles = [], []
    for k in range(5):
        mesh = rng.random((nb, n, n), dtype=np.float32)
        xp = mesh.sum(axis=2)
        at = f.All if k % 2 == 0 else f.Defaults
        f.append(k * 0.1, at, mesh=mesh, xproj=xp, yproj=mesh.sum(axis=1),
                 length=[1.0, 1.1], position=[0.0, 0.1], spread=[1, 1],
                 average=[0, 0], population=[0.6, 0.4])
        f.append_field([k, -k], rng.random((nb, 32), dtype=np.float32))
        f.append_wake(np.full((nb, n), k, np.float32))
        f.appendRFKicks(np.array([[0.1 * k, 1.0]] * (k % 3), np.float32))
        f.appendTracks(np.ones((3, 2), np.float32) * k)
        profiles.append(xp)
        if at == f.All:
            meshes.append(mesh)
    f.append(0.5, f.PhaseSpace, mesh=meshes[-1])
    meshes.append(meshes[-1])
    assert f.rows("/PhaseSpace/data") == 4
    assert f.rows("/Info/AxisValues_t") == 5
    assert f.rows("/RFKicks/data") == 0 + 1 + 2 + 0 + 1
    f.close()
    # readPhaseSpace: positive and negative step numbers (:452)
    for step in (0, 1, 3, -1, -4):
        ps, steps = H.readPhaseSpace(path, step)
        assert steps == 4
        assert np.array_equal(ps, meshes[step])
    got = _read(path, "/BunchProfile/data", str(tmp_path))
    assert np.array_equal(got.reshape(5, nb, n), np.array(profiles))
    t = _read(path, "/PhaseSpace/axis0", str(tmp_path))
    assert np.allclose(t, [0.0, 0.2, 0.4, 0.5])
    assert np.array_equal(_read(path, "/CSR/Intensity/data",
                                str(tmp_path)).reshape(5, 2)[:, 1],
                          -np.arange(5, dtype=np.float32))
    with pytest.raises(RuntimeError):
        H.readPhaseSpace(str(tmp_path / "missing.h5"))


def test_library_exports_every_declared_symbol(H):
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    text = open(os.path.join(root, "include", "vfps_hdf5.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    syms = sorted(set(re.findall(r"VFPS_H5_API\s+[\w\s\*]+?\b(vfps_h5_\w+)\s*\(",
                                 text)))
    assert len(syms) >= 14
    assert syms == H.exported_symbols()
    for s in syms:
        assert hasattr(H.lib(), s), s
