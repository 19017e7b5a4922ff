"""Result and checkpoint files in the layout of Inovesa's HDF5File
(inc/IO/HDF5File.hpp, src/IO/HDF5File.cpp) through libvfps_h5.so, resume
from /PhaseSpace/data (HDF5File::readPhaseSpace), and a writer that runs a
Step with the output steps of src/main.cpp:961-997.

The library is host only: the step's quantities are read from the device
(Step.moments, the CSR and wake fields) and handed over as host arrays.
"""
import ctypes as C
import os

import numpy as np

from .vfps import ElectricField, PhaseSpace

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

C_LIGHT = 2.99792458e8
FACTOR4OHMS = 1.0  # Impedance::factor4Ohms (inc/Z/Impedance.hpp:91)

_u32, _i32, _f32, _f64, _vp, _int = (C.c_uint32, C.c_int32, C.c_float,
                                     C.c_double, C.c_void_p, C.c_int)


class Info(C.Structure):
    """vfps_h5_info (include/vfps_hdf5.h)."""
    _fields_ = [("nb", _u32), ("nx", _u32), ("ny", _u32), ("nbuckets", _u32),
                ("buckets", _vp), ("maxn", _u32), ("imp_size", _u32),
                ("nparticles", _u32), ("axis_z", _vp), ("axis_f", _vp),
                ("meter", _f64), ("electronvolt", _f64), ("hertz", _f64),
                ("t_sync", _f64), ("f_rev", _f64), ("current", _f64),
                ("charge", _f64), ("volts", _f64),
                ("factor4WattPerHertz", _f64), ("factor4Watts", _f64),
                ("factor4Ohms", _f64), ("imp_real", _vp), ("imp_imag", _vp),
                ("version", _i32 * 3), ("build", C.c_char_p)]


_SIGS = {
    "vfps_h5_last_error": (C.c_char_p, []),
    "vfps_h5_create": (_int, [C.c_char_p, C.POINTER(Info), _vp]),
    "vfps_h5_close": (_int, [_vp]),
    "vfps_h5_append_ps": (_int, [_vp, _f32, _int, _vp, _vp, _vp, _vp, _vp,
                                 _vp, _vp, _vp]),
    "vfps_h5_append_csr": (_int, [_vp, _vp, _vp]),
    "vfps_h5_append_padded": (_int, [_vp, _vp, _vp]),
    "vfps_h5_append_rfkicks": (_int, [_vp, _vp, _u32]),
    "vfps_h5_append_tracks": (_int, [_vp, _vp]),
    "vfps_h5_append_wake": (_int, [_vp, _vp]),
    "vfps_h5_add_parameter_f64": (_int, [_vp, C.c_char_p, C.c_char_p, _f64]),
    "vfps_h5_add_parameter_i64": (_int, [_vp, C.c_char_p, C.c_char_p,
                                         C.c_int64]),
    "vfps_h5_add_parameter_str": (_int, [_vp, C.c_char_p, C.c_char_p,
                                         C.c_char_p]),
    "vfps_h5_add_parameter_typed": (_int, [_vp, C.c_char_p, C.c_char_p,
                                           _int, _vp]),
    "vfps_h5_rows": (_int, [_vp, C.c_char_p, C.POINTER(C.c_uint64)]),
    "vfps_h5_read_phasespace": (_int, [C.c_char_p, C.c_int64,
                                       C.POINTER(_u32), C.POINTER(_u32),
                                       C.POINTER(C.c_uint64), _vp]),
}


def lib():
    """libvfps_h5.so (built by inovesa_amd/build.py build_h5)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "lib", "libvfps_h5.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with "
                              "inovesa_amd/build.py (needs HDF5)")
        L = C.CDLL(path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def exported_symbols():
    return sorted(_SIGS)


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().vfps_h5_last_error().decode())


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


def _f(a, shape=None):
    a = np.ascontiguousarray(a, np.float32)
    return a if shape is None else a.reshape(shape)


class HDF5File:
    """HDF5File(filename, ps, ef, imp, nparticles, t_sync, f_rev)
    (src/IO/HDF5File.cpp:29-330), with the constructor's inputs given as
    values: nb bunches of n x n meshes, the position axis axis_z and its
    "Meter" scale, the energy scale "ElectronVolt", the field (bucket
    numbers, maxn = nmax/2, frequency axis and its "Hertz" scale, volts,
    factor4WattPerHertz, factor4Watts) and the impedance table."""

    All, Defaults, PhaseSpace = 0, 1, 2   # HDF5File::AppendType

    def __init__(self, filename, nb, n, axis_z, meter, electronvolt, t_sync,
                 f_rev, current, charge, buckets=(), maxn=0, axis_f=None,
                 hertz=0.0, volts=0.0, factor4WattPerHertz=0.0,
                 factor4Watts=0.0, impedance=None, nparticles=0,
                 version=(1, 1, 0), build="inovesa_amd (MI355X)"):
        self.nb, self.n, self.maxn = nb, n, maxn
        self.nparticles = nparticles
        self._axis_z = _f(axis_z)
        self._axis_f = _f(axis_f) if axis_f is not None else None
        self._buckets = np.ascontiguousarray(buckets, np.uint32)
        imp = None if impedance is None else np.asarray(impedance)
        self.imp_size = 0 if imp is None else len(imp) // 2
        self._ir = None if imp is None else _f(imp.real[:self.imp_size])
        self._ii = None if imp is None else _f(imp.imag[:self.imp_size])
        self._build = build.encode()
        info = Info(nb=nb, nx=n, ny=n, nbuckets=self._buckets.size,
                    buckets=_p(self._buckets) if self._buckets.size else None,
                    maxn=maxn, imp_size=self.imp_size, nparticles=nparticles,
                    axis_z=_p(self._axis_z), axis_f=_p(self._axis_f),
                    meter=meter, electronvolt=electronvolt, hertz=hertz,
                    t_sync=t_sync, f_rev=f_rev, current=current,
                    charge=charge, volts=volts,
                    factor4WattPerHertz=factor4WattPerHertz,
                    factor4Watts=factor4Watts, factor4Ohms=FACTOR4OHMS,
                    imp_real=_p(self._ir), imp_imag=_p(self._ii),
                    version=(_i32 * 3)(*version), build=self._build)
        h = C.c_void_p()
        _check(lib().vfps_h5_create(os.fsencode(filename), C.byref(info),
                                    C.byref(h)))
        self.h = h
        self.filename = filename

    def append(self, t, at=1, mesh=None, xproj=None, yproj=None, length=None,
               position=None, spread=None, average=None, population=None):
        """append(const PhaseSpace&, t, at) (src/IO/HDF5File.cpp:391-419)."""
        nb, n = self.nb, self.n
        arr = lambda a, s: None if a is None else _f(a, s)
        m = arr(mesh, (nb, n, n))
        xs = [arr(xproj, (nb, n)), arr(yproj, (nb, n))]
        vs = [arr(v, (nb,)) for v in (length, position, spread, average,
                                      population)]
        _check(lib().vfps_h5_append_ps(self.h, float(t), int(at), _p(m),
                                       *[_p(x) for x in xs],
                                       *[_p(v) for v in vs]))

    def append_field(self, power, spectrum=None):
        """append(const ElectricField*, fullspectrum) (:341-347)."""
        p = _f(power, (self.nb,))
        s = None if spectrum is None else _f(spectrum, (self.nb, self.maxn))
        _check(lib().vfps_h5_append_csr(self.h, _p(p), _p(s)))

    def appendPadded(self, profile, wake):
        p = _f(profile)[:self.imp_size]
        w = _f(wake)[:self.imp_size]
        _check(lib().vfps_h5_append_padded(self.h, _p(_f(p)), _p(_f(w))))

    def appendRFKicks(self, kicks):
        k = _f(kicks).reshape(-1, 2)
        _check(lib().vfps_h5_append_rfkicks(self.h, _p(k), k.shape[0]))

    def appendTracks(self, positions):
        p = _f(positions, (self.nparticles, 2))
        _check(lib().vfps_h5_append_tracks(self.h, _p(p)))

    def append_wake(self, wake):
        """append(const WakeKickMap*) (:421-424): the wake potential."""
        _check(lib().vfps_h5_append_wake(self.h, _p(_f(wake,
                                                       (self.nb, self.n)))))

    def addParameterToGroup(self, group, name, value):
        g, k = group.encode(), name.encode()
        if isinstance(value, str):
            rc = lib().vfps_h5_add_parameter_str(self.h, g, k, value.encode())
        elif isinstance(value, (bool, int, np.integer)):
            rc = lib().vfps_h5_add_parameter_i64(self.h, g, k, int(value))
        else:
            rc = lib().vfps_h5_add_parameter_f64(self.h, g, k, float(value))
        _check(rc)

    def addTypedParameter(self, group, name, ctype, value):
        """An attribute of the HDF5 type of ctype (float, double, uint32 or
        int32), as ProgramOptions::save(HDF5File*) writes the options."""
        kinds = {"float": (0, C.c_float), "double": (1, C.c_double),
                 "uint32": (2, C.c_uint32), "int32": (3, C.c_int32)}
        kind, ct = kinds[ctype]
        v = ct(value)
        _check(lib().vfps_h5_add_parameter_typed(self.h, group.encode(),
                                                 name.encode(), kind,
                                                 C.byref(v)))

    def rows(self, dataset):
        r = C.c_uint64()
        _check(lib().vfps_h5_rows(self.h, dataset.encode(), C.byref(r)))
        return r.value

    def close(self):
        if getattr(self, "h", None):
            lib().vfps_h5_close(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def readPhaseSpace(filename, step=-1):
    """HDF5File::readPhaseSpace (src/IO/HDF5File.cpp:428-495): phase space
    number `step` of /PhaseSpace/data (negative: from the end), as
    [nb][n][n] float32, and the number of phase spaces in the file."""
    nb, n, steps = _u32(), _u32(), C.c_uint64()
    path = os.fsencode(filename)
    _check(lib().vfps_h5_read_phasespace(path, step, C.byref(nb), C.byref(n),
                                         C.byref(steps), None))
    out = np.empty((nb.value, n.value, n.value), np.float32)
    _check(lib().vfps_h5_read_phasespace(path, step, None, None, None,
                                         _p(out)))
    return out, steps.value


def ruler(n, vmin, vmax):
    """inc/PS/Ruler.hpp:31-51 in float32: values and delta."""
    vmin, vmax = np.float32(vmin), np.float32(vmax)
    delta = (vmax - vmin) / np.float32(n - 1)
    return vmin + np.arange(n, dtype=np.float32) * delta, np.float32(delta)


class StepWriter:
    """The output of src/main.cpp:908-997 for a Step: at every outstep-th
    step (simulationstep % outstep == 0, before the maps of that step) the
    phase space's projections and moments, the CSR power and spectrum of
    the radiation field (updateCSR(fc)) and the wake potential the step
    kicks with; the phase space itself every h5save-th output (0: only the
    initial one), padded profile and wake at the start.

    The CSR spectrum and the wake potential come from ElectricFields over a
    side phase space that holds the step's current profile (the fused step
    forms the wake of a profile inside the next step)."""

    def __init__(self, filename, step, desc, derived, zwake, zcsr, outstep,
                 h5save=0, dev=None):
        d = derived
        ring = d["ring"]
        n = desc["n"]
        self.step, self.outstep, self.h5save = step, outstep, h5save
        self.steps_per_ts = d["steps"]
        self.fc = np.float32(desc.get("fc", ring["cutoff_frequency"]))
        self.simulationstep = 0
        self.outstepnr = 0
        axis_z, dq = ruler(n, desc["qmin"], desc["qmax"])
        _, dp = ruler(n, desc["pmin"], desc["pmax"])
        meter = float(np.float32(desc["qscale"]))
        ev = float(np.float32(desc["pscale"]))
        f_rev = float(np.float32(ring["f_rev"]))
        nmax = zwake.n if zwake is not None else zcsr.n
        faxis, _ = ruler(nmax, 0.0, np.float32(1.0) / dq)
        hertz = C_LIGHT / meter
        Ib = d["Ib"]
        wph = 2 * FACTOR4OHMS * Ib * Ib / f_rev
        self.side = PhaseSpace(n, desc["qmin"], desc["qmax"], desc["qscale"],
                               desc["pmin"], desc["pmax"], desc["pscale"],
                               dev=dev)
        rp = np.float32(desc["revolutionpart"])
        self.rad = ElectricField(self.side, zcsr, [0], 0, f_rev, rp)
        self.wake = (ElectricField(self.side, zwake, [0],
                                   desc["spacing_bins"], f_rev, rp,
                                   desc["wakescaling"])
                     if zwake is not None and desc.get("use_wake", 1)
                     else None)
        self.file = HDF5File(
            filename, 1, n, axis_z, meter, ev, 1.0 / d["f_s"], f_rev, Ib,
            Ib / f_rev, buckets=[0], maxn=nmax // 2,
            axis_f=faxis[:nmax // 2], hertz=hertz,
            volts=float(dp) * ev / float(rp), factor4WattPerHertz=wph,
            factor4Watts=wph * hertz,
            impedance=zwake.data() if zwake is not None else None)
        self.file.addParameterToGroup("/Info", "ShieldingParameter",
                                      float(d["shield"]))
        # initial conditions (src/main.cpp:912-922)
        m = step.moments()
        self.side.setProjection(0, m["xproj"])
        if self.wake is not None:
            self.wake.wakePotential()
            self.file.appendPadded(self.wake.getPaddedBunchProfiles(),
                                   self._padded_wake())
        if h5save == 0:
            mesh, _, _ = step.mesh()
            self.file.append(0.0, HDF5File.PhaseSpace, mesh=mesh)

    def _padded_wake(self):
        """getPaddedWakePotential() of the last wakePotential()."""
        return self.wake.getPaddedWakePotential()

    def _output(self):
        st, f = self.step, self.file
        m = st.moments()
        at = (HDF5File.All if self.h5save > 0 and
              self.outstepnr % self.h5save == 0 else HDF5File.Defaults)
        mesh = st.mesh()[0] if at == HDF5File.All else None
        f.append(self.simulationstep / self.steps_per_ts, at, mesh=mesh,
                 xproj=m["xproj"], yproj=m["yproj"], length=m["length"],
                 position=m["position"], spread=m["spread"],
                 average=m["average"], population=m["population"])
        self.side.setProjection(0, m["xproj"])
        power, spec = self.rad.updateCSR(self.fc)
        f.append_field(power, spec[:, :f.maxn])
        if self.wake is not None:
            f.append_wake(self.wake.wakePotential())
        self.outstepnr += 1

    def run(self, nsteps):
        """nsteps passes of the loop body, writing at the output steps."""
        end = self.simulationstep + nsteps
        while self.simulationstep < end:
            if self.outstep > 0 and self.simulationstep % self.outstep == 0:
                self._output()
            k = end - self.simulationstep
            if self.outstep > 0:
                k = min(k, self.outstep - self.simulationstep % self.outstep)
            self.step.run(k)
            self.simulationstep += k

    def finish(self):
        """After the loop (src/main.cpp:1079-1099): the final phase space
        with everything, the fields and the padded arrays."""
        st, f = self.step, self.file
        m = st.moments()
        mesh = st.mesh()[0]
        f.append(self.simulationstep / self.steps_per_ts, HDF5File.All,
                 mesh=mesh, xproj=m["xproj"], yproj=m["yproj"],
                 length=m["length"], position=m["position"],
                 spread=m["spread"], average=m["average"],
                 population=m["population"])
        self.side.setProjection(0, m["xproj"])
        power, spec = self.rad.updateCSR(self.fc)
        f.append_field(power, spec[:, :f.maxn])
        if self.wake is not None:
            f.append_wake(self.wake.wakePotential())
            f.appendPadded(self.wake.getPaddedBunchProfiles(),
                           self._padded_wake())
        f.close()
