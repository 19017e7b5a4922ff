"""The reference's program options (src/IO/ProgramOptions.cpp) and the
setup derived from them (src/main.cpp:163-355), on the CPU."""
import math

import numpy as np
import pytest

from inovesa_amd.options import OptionError, ProgramOptions

CFG = """# a configuration in the reference's format
GridSize=128
BunchCurrent=1e-3
BunchCurrent=0
BunchCurrent=5e-4
RFVoltage=1.2e6          # alias of AcceleratingVoltage
SyncFreq = 3e4           # alias of SynchrotronFrequency
steps=800                # alias of StepsPerTs
VacuumGap=-1
HaissinskiIterations=3   # compatibility: ignored
RotationType=2
LinearRF=false
InterpolateClamped=1
output=/dev/null
"""


def test_defaults_are_the_reference_defaults():
    """src/IO/ProgramOptions.cpp:28-282 and Appendix of the defaults:
    E0 1.3 GeV, sigma_delta 4.7e-4, f_rev 9 MHz, H 50, V 1 MV, alpha0 4e-3,
    G 3 cm, PhaseSpaceSize 12, GridSize 256, padding 8 (rounded), 1000
    steps per period, 5 periods, 3 mA, outstep 100, cubic interpolation and
    derivative, full Fokker-Planck with stochastic tracking."""
    o = ProgramOptions()
    exp = dict(BeamEnergy=1.3e9, BeamEnergySpread=4.7e-4,
               RevolutionFrequency=9e6, HarmonicNumber=50.0,
               AcceleratingVoltage=1e6, alpha0=float(np.float32(4e-3)),
               VacuumGap=0.03, PhaseSpaceSize=12.0, GridSize=256,
               padding=8.0, RoundPadding=True, StepsPerTs=1000,
               rotations=5.0, outstep=100, InterpolationPoints=4,
               derivation=4, FPType=3, FPTrack=3, RenormalizeCharge=0,
               CutoffFreq=float(np.float32(23e9)), LinearRF=True,
               UseCSR=True, DampingTime=-1.0, BendingRadius=-1.0)
    for k, v in exp.items():
        assert o[k] == v, k
    assert o["BunchCurrent"] == [float(np.float32(3e-3))]


def test_defaults_derive_the_documented_ring():
    from inovesa_amd import derive
    a = derive(**ProgramOptions().ring())
    b = derive()
    for k in ("qscale", "V_eff", "f_s", "Ith", "padded_bins", "spacing_bins",
              "wakescaling", "e1", "laststep"):
        assert a[k] == pytest.approx(b[k], rel=1e-7), k
    assert a["laststep"] == 5000


def test_config_file_aliases_and_multivalued(tmp_path):
    p = tmp_path / "run.cfg"
    p.write_text(CFG)
    o = ProgramOptions().parse(["-c", str(p), "--GridSize", "64",
                                "-T", "0.5"])
    assert o["GridSize"] == 64            # the command line wins
    assert o["rotations"] == 0.5
    assert o["AcceleratingVoltage"] == 1.2e6
    assert o["SynchrotronFrequency"] == float(np.float32(3e4))
    assert o["StepsPerTs"] == 800
    assert o["LinearRF"] is False and o["InterpolateClamped"] is True
    assert o["output"] == ""              # /dev/null: no output
    assert o["BunchCurrent"] == [float(np.float32(x)) for x in (1e-3, 0, 5e-4)]
    # buckets count from the end of the pattern (src/main.cpp:259-266)
    buckets, currents = o.bunches()
    assert buckets == [2, 0] and len(currents) == 2
    ring = o.ring()
    assert ring["bunch_current"] == pytest.approx(1.5e-3, rel=1e-6)
    with pytest.raises(OptionError):
        ProgramOptions().read_cfg("NoSuchOption=1\n")
    with pytest.raises(OptionError):
        ProgramOptions().read_cfg("GridSize=abc\n")


def test_save_and_reload_round_trip(tmp_path):
    o = ProgramOptions().parse(["-s", "512", "-I", "1e-3", "2e-3",
                                "--VacuumGap", "0.032", "--LinearRF", "0",
                                "--outstep", "50", "--Impedance", "z.dat"])
    text = o.save_text()
    assert "run_anyway" not in text
    assert text.count("BunchCurrent=") == 2
    r = ProgramOptions()
    r.read_cfg(text)
    for k in ("GridSize", "BunchCurrent", "VacuumGap", "LinearRF", "outstep",
              "Impedance", "alpha0", "StepsPerTs", "CutoffFreq"):
        assert r[k] == o[k], k


def test_steps_per_revolution_overrides_steps_per_period():
    """src/main.cpp:284-286: StepsPerRevolution > 0 sets
    steps = StepsPerRevolution * f_rev / f_s."""
    from inovesa_amd import derive
    o = ProgramOptions().parse(["--StepsPerRevolution", "2"])
    d = derive(**o.ring())
    steps = 2 * 9e6 / d["f_s"]
    assert d["steps"] == pytest.approx(steps, rel=1e-12)
    assert d["revolutionpart"] == pytest.approx(0.5, rel=1e-9)
    assert d["angle"] == pytest.approx(2 * math.pi / steps, rel=1e-6)


def test_impedance_file_reader(tmp_path):
    """Impedance::readData (src/Z/Impedance.cpp:93-111) on the reference's
    test/impedance.dat (index, real, imaginary: n-2 and 2.5-n), the last
    line read twice at the end of the stream counted once."""
    from inovesa_amd.run import read_impedance_file
    p = tmp_path / "impedance.dat"
    rows = [f"{i}\t{i - 2}\t {2.5 - i}" for i in range(1, 7)]
    p.write_text("\n".join(rows + [rows[-1]]) + "\n")
    z = read_impedance_file(str(p), 8)
    np.testing.assert_array_equal(z[:6].real, np.arange(1, 7) - 2)
    np.testing.assert_array_equal(z[:6].imag, 2.5 - np.arange(1, 7))
    assert np.all(z[6:] == 0)
