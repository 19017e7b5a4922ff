"""The reference's program options (src/IO/ProgramOptions.cpp:13-305,
inc/IO/ProgramOptions.hpp): the same option names, short forms, types and
defaults, read from configuration files in the reference's format
(`name=value` lines, `#` comments, repeated lines adding to multi-valued
options), from the command line, and saved back as the reference saves
them (ProgramOptions::save, :384-490), so that a reference configuration
runs unchanged.
"""
import math
import shlex

import numpy as np

# name: (type, default, short option, group); types are the reference's
# member types: "float" (meshaxis_t / frequency_t), "double", "uint32",
# "int32", "int64", "bool", "str", "floatvec" (std::vector<integral_t>)
OPTIONS = {
    # physical parameters (:28-138)
    "alpha0": ("float", 4e-3, None, "phys"),
    "alpha1": ("float", 0.0, None, "phys"),
    "alpha2": ("float", 0.0, None, "phys"),
    "SynchrotronFrequency": ("float", 0.0, "f", "phys"),
    "RevolutionFrequency": ("float", 9e6, "F", "phys"),
    "DampingTime": ("double", -1.0, "d", "phys"),
    "HarmonicNumber": ("float", 50.0, "H", "phys"),
    "InitialDistFile": ("str", "", "i", "phys"),
    "InitialDistStep": ("int64", -1, None, "phys"),
    "InitialDistZoom": ("double", 1.0, None, "phys"),
    "BunchCurrent": ("floatvec", [3e-3], "I", "phys"),
    "BendingRadius": ("double", -1.0, "R", "phys"),
    "BeamEnergy": ("double", 1.3e9, "E", "phys"),
    "BeamEnergySpread": ("double", 4.7e-4, "e", "phys"),
    "Impedance": ("str", "", "Z", "phys"),
    "VacuumGap": ("double", 0.03, "G", "phys"),
    "UseCSR": ("bool", True, None, "phys"),
    "CollimatorRadius": ("double", 0.0, None, "phys"),
    "WallConductivity": ("double", 0.0, None, "phys"),
    "WallSusceptibility": ("double", 0.0, None, "phys"),
    "CutoffFreq": ("float", 23e9, None, "phys"),
    "AcceleratingVoltage": ("double", 1e6, "V", "phys"),
    "LinearRF": ("bool", True, None, "phys"),
    "RFAmplitudeSpread": ("double", 0.0, None, "phys"),
    "RFPhaseSpread": ("double", 0.0, None, "phys"),
    "RFPhaseModAmplitude": ("double", 0.0, None, "phys"),
    "RFPhaseModFrequency": ("double", 0.0, None, "phys"),
    # program parameters (:139-175); cldev and ForceOpenGLVersion are read
    # and ignored (no OpenCL, no display)
    "cldev": ("int32", 0, None, "prog"),
    "ForceOpenGLVersion": ("int32", 2, None, "prog"),
    "gui": ("bool", False, "g", "prog"),
    "output": ("str", "", "o", "prog"),
    "outstep": ("uint32", 100, "n", "prog"),
    "SavePhaseSpace": ("uint32", 0, None, "prog"),
    "tracking": ("str", "", None, "prog"),
    "verbose": ("bool", False, "v", "prog"),
    "run_anyway": ("bool", False, None, "prog"),
    # simulation parameters (:218-282)
    "StepsPerTs": ("uint32", 1000, "N", "simul"),
    "StepsPerRevolution": ("double", 0.0, None, "simul"),
    "padding": ("double", 8.0, "p", "simul"),
    "RoundPadding": ("bool", True, None, "simul"),
    "PhaseSpaceSize": ("float", 12.0, "P", "simul"),
    "PhaseSpaceShiftX": ("float", 0.0, None, "simul"),
    "PhaseSpaceShiftY": ("float", 0.0, None, "simul"),
    "RenormalizeCharge": ("int32", 0, None, "simul"),
    "FPType": ("uint32", 3, None, "simul"),
    "FPTrack": ("uint32", 3, None, "simul"),
    "GridSize": ("uint32", 256, "s", "simul"),
    "rotations": ("double", 5.0, "T", "simul"),
    "derivation": ("uint32", 4, None, "simul"),
    "InterpolationPoints": ("uint32", 4, None, "simul"),
    "InterpolateClamped": ("bool", False, None, "simul"),
}

# compatibility options (:283-304): read and ignored, or aliases
IGNORED = {"HaissinskiIterations", "InitialDistParam", "RotationType",
           "SaveSourceMap"}
ALIASES = {"RFVoltage": "AcceleratingVoltage",
           "SyncFreq": "SynchrotronFrequency", "steps": "StepsPerTs"}
SHORT = {v[2]: k for k, v in OPTIONS.items() if v[2]}
SHORT["c"] = "config"


class OptionError(ValueError):
    pass


def _is_option(arg):
    """An argument that starts an option (not a negative number)."""
    if not arg.startswith("-") or len(arg) < 2:
        return False
    try:
        float(arg)
        return False
    except ValueError:
        return True


def _bool(text):
    t = text.strip().lower()
    if t in ("1", "true", "yes", "on"):
        return True
    if t in ("0", "false", "no", "off"):
        return False
    raise OptionError(f"not a boolean: {text!r}")


def convert(name, text):
    """A value of option `name` from its text, in the option's type."""
    typ = OPTIONS[name][0]
    try:
        if typ == "float":
            return float(np.float32(float(text)))
        if typ == "double":
            return float(text)
        if typ in ("uint32", "int32", "int64"):
            v = int(float(text)) if "e" in text.lower() else int(text, 0)
            if typ == "uint32" and not 0 <= v < 2 ** 32:
                raise OptionError(f"{name}: {v} out of range")
            return v
        if typ == "bool":
            return _bool(text)
        if typ == "floatvec":
            return [float(np.float32(float(t))) for t in text.split()]
        return text.strip()
    except ValueError as e:
        raise OptionError(f"{name}: cannot read {text!r} ({e})") from None


class ProgramOptions:
    """vfps::ProgramOptions: defaults, configuration file, command line."""

    def __init__(self):
        def default(typ, v):
            # float members hold their defaults as float
            if typ == "float":
                return float(np.float32(v))
            if typ == "floatvec":
                return [float(np.float32(x)) for x in v]
            return v
        self.values = {k: default(v[0], v[1]) for k, v in OPTIONS.items()}
        self.given = set()        # options set explicitly (_vm entries)
        self.configfile = "default.cfg"

    # -- reading ----------------------------------------------------------
    def set(self, name, value, from_text=True):
        name = ALIASES.get(name, name)
        if name in IGNORED:
            return
        if name not in OPTIONS:
            raise OptionError(f"unknown option {name!r}")
        v = convert(name, value) if from_text else value
        if OPTIONS[name][0] == "floatvec" and from_text:
            # a multi-valued option collects every occurrence
            if name in self.given:
                self.values[name] = self.values[name] + v
            else:
                self.values[name] = v
        else:
            self.values[name] = v
        self.given.add(name)

    def read_cfg(self, text):
        """po::parse_config_file: name=value lines; [section] headers
        prefix the names (the reference's files have none)."""
        section = ""
        for lineno, raw in enumerate(text.splitlines(), 1):
            line = raw.split("#", 1)[0].strip()
            if not line:
                continue
            if line.startswith("[") and line.endswith("]"):
                section = line[1:-1].strip() + "."
                continue
            if "=" not in line:
                raise OptionError(f"line {lineno}: expected name=value")
            k, v = line.split("=", 1)
            self.set(section + k.strip(), v.strip())

    def load(self, path):
        with open(path) as f:
            self.read_cfg(f.read())
        self.configfile = path

    def parse(self, argv):
        """The command line (:313-377): --name value, --name=value, -x
        value; --config/-c loads a configuration file first, whose values
        the command line then overrides (the reference stores the command
        line first, and po::store keeps the first value of an option)."""
        cli = []
        cfg = None
        i = 0
        argv = list(argv)
        while i < len(argv):
            a = argv[i]
            if a.startswith("--"):
                name, _, val = a[2:].partition("=")
            elif a.startswith("-") and len(a) >= 2:
                name, val = SHORT.get(a[1], a[1]), a[2:]
            else:
                raise OptionError(f"unexpected argument {a!r}")
            name = ALIASES.get(name, name)
            typ = OPTIONS.get(name, ("str",))[0]
            if not val:
                implicit = typ == "bool" and (i + 1 >= len(argv) or
                                              argv[i + 1].startswith("-"))
                if implicit:          # gui, verbose, run_anyway
                    val = "true"
                elif typ == "floatvec":
                    vals = []
                    while i + 1 < len(argv) and not _is_option(argv[i + 1]):
                        i += 1
                        vals.append(argv[i])
                    val = " ".join(vals)
                else:
                    i += 1
                    if i >= len(argv):
                        raise OptionError(f"{name}: value missing")
                    val = argv[i]
            if name == "config":
                cfg = val
            else:
                cli.append((name, val))
            i += 1
        if cfg == "/dev/null":
            cfg = None
        if cfg:
            self.load(cfg)
        for name, val in cli:
            if OPTIONS.get(name, ("",))[0] == "floatvec":
                self.given.discard(name)
            self.set(name, val)
        for key in ("output", "InitialDistFile", "tracking"):
            if self.values[key] == "/dev/null":
                self.values[key] = ""
        return self

    # -- access ------------------------------------------------------------
    def __getitem__(self, name):
        return self.values[ALIASES.get(name, name)]

    def __setitem__(self, name, value):
        self.set(name, value, from_text=isinstance(value, str))

    # -- saving -------------------------------------------------------------
    def _fmt(self, name):
        typ = OPTIONS[name][0]
        v = self.values[name]
        if typ == "bool":
            return "1" if v else "0"
        if typ in ("float", "double"):
            # std::ostream's default: 6 significant digits
            return f"{v:.6g}"
        if typ == "floatvec":
            return " ".join(f"{x:.6g}" for x in v)
        return str(v)

    def save_text(self, version="inovesa_amd"):
        """ProgramOptions::save(std::string) (:384-450): every option with
        a value, name=value, in name order, without run_anyway.  (The
        reference writes alpha0=0 whenever SynchrotronFrequency is 0, the
        default, which a reload would then take literally; alpha0 keeps its
        value here.)"""
        lines = [f"# {version}"]
        for name in sorted(OPTIONS, key=str):
            if name == "run_anyway":
                continue
            if OPTIONS[name][0] == "str" and not self.values[name]:
                continue
            if OPTIONS[name][0] == "floatvec":
                for x in self.values[name]:
                    lines.append(f"{name}={x:.6g}")
                continue
            lines.append(f"{name}={self._fmt(name)}")
        return "\n".join(lines) + "\n"

    def save(self, path):
        with open(path, "w") as f:
            f.write(self.save_text())

    def save_h5(self, h5file):
        """ProgramOptions::save(HDF5File*) (:460-490): every numeric option
        as an attribute of /Info/Parameters in its own type."""
        kinds = {"float": "float", "double": "double", "uint32": "uint32",
                 "int32": "int32"}
        for name, (typ, *_rest) in sorted(OPTIONS.items()):
            if typ in kinds:
                h5file.addTypedParameter("/Info/Parameters", name, kinds[typ],
                                         self.values[name])

    # -- derived quantities of src/main.cpp ---------------------------------
    def ring(self):
        """The ring parameters of vfps_derive (src/main.cpp:163-355) for
        the filled buckets' total current."""
        v = self.values
        return dict(grid_size=v["GridSize"], pq_size=v["PhaseSpaceSize"],
                    shift_x=v["PhaseSpaceShiftX"],
                    shift_y=v["PhaseSpaceShiftY"],
                    energy_spread=v["BeamEnergySpread"],
                    beam_energy=v["BeamEnergy"],
                    bending_radius=v["BendingRadius"],
                    f_rev=v["RevolutionFrequency"],
                    harmonic_number=v["HarmonicNumber"],
                    rf_voltage=v["AcceleratingVoltage"],
                    f_s=v["SynchrotronFrequency"], alpha0=v["alpha0"],
                    alpha1=v["alpha1"], alpha2=v["alpha2"],
                    vacuum_gap=v["VacuumGap"], damping_time=v["DampingTime"],
                    padding=v["padding"],
                    round_padding=int(v["RoundPadding"]),
                    steps_per_ts=v["StepsPerTs"],
                    rotations=v["rotations"],
                    bunch_current=float(sum(self.bunches()[1])),
                    cutoff_frequency=v["CutoffFreq"],
                    steps_per_trev=v["StepsPerRevolution"])

    def bunches(self):
        """src/main.cpp:254-273: (bucket numbers, currents) of the filled
        buckets; buckets count from the end of the filling pattern."""
        fill = self.values["BunchCurrent"]
        buckets, currents = [], []
        for i, c in enumerate(fill):
            if c > 0:
                buckets.append(len(fill) - 1 - i)
                currents.append(float(np.float32(c)))
        return buckets, currents

    def rf_noise(self, dt):
        """src/main.cpp:330-345: RF phase noise and modulation amplitude in
        rad, amplitude noise, modulation step per time step."""
        v = self.values
        two_pi = 2 * math.pi
        return (max(0.0, v["RFPhaseSpread"] / 360.0 * two_pi),
                max(0.0, v["RFAmplitudeSpread"]),
                max(0.0, v["RFPhaseModAmplitude"] / 360.0 * two_pi),
                v["RFPhaseModFrequency"] * dt)


def parse_args(argv):
    return ProgramOptions().parse(argv)


def split_cli(text):
    return shlex.split(text)
