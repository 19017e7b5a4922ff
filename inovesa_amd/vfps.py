"""Inovesa's phase-space interfaces over the MI355X C ABI.

Class and method names follow the reference (inc/PS/PhaseSpace.hpp,
inc/SM/*.hpp, inc/PS/ElectricField.hpp, inc/Z/*.hpp) so that code written
against them reads the same; every method forwards to libvfps.so and runs on
the GPU.  Arguments the reference rejects with std::invalid_argument raise
VfpsInvalid (a ValueError).
"""
import ctypes as C

import numpy as np

from ._lib import (Derived, RingParams, StepDesc, VfpsError, VfpsInvalid,
                   call, lib)

__all__ = ["Device", "PhaseSpace", "KickMap", "DriftMap", "RFKickMap",
           "RotationMap", "FokkerPlanckMap", "Identity", "WakePotentialMap",
           "Impedance", "FreeSpaceCSR", "ParallelPlatesCSR", "ResistiveWall",
           "ConstImpedance", "CollimatorImpedance", "makeImpedance",
           "ElectricField", "Step", "derive", "airy", "VfpsError",
           "VfpsInvalid", "InterpolationType", "Axis", "FPType",
           "DerivationType", "toeplitz_kernels", "Batch", "MLP",
           "gemm_bf16", "round_bf16", "wake_potential_noalloc",
           "DynamicRFKickMap", "dynrf_modulation"]


class InterpolationType:  # inc/SM/SourceMap.hpp:22-27
    none = 1
    linear = 2
    quadratic = 3
    cubic = 4


class Axis:  # inc/SM/KickMap.hpp:29-31
    x = 0
    y = 1


class FPType:  # inc/SM/FokkerPlanckMap.hpp:31-36
    none = 0
    damping_only = 1
    diffusion_only = 2
    full = 3


class DerivationType:  # inc/SM/FokkerPlanckMap.hpp:43-46 (+ five_point)
    two_sided = 3
    cubic = 4
    five_point = 5


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Device:
    """The device handle that replaces the reference's OpenCL handler."""

    _default = None

    def __init__(self, ordinal=0):
        h = C.c_void_p()
        call("vfps_dev_open", ordinal, C.byref(h))
        self.h = h

    @classmethod
    def default(cls):
        if cls._default is None:
            cls._default = Device(0)
        return cls._default

    def sync(self):
        call("vfps_sync", self.h)

    def close(self):
        if self.h:
            lib.vfps_dev_close(self.h)
            self.h = None


class PhaseSpace:
    """inc/PS/PhaseSpace.hpp:40-552 on the device; data layout [nb][n][n]."""

    def __init__(self, n, qmin, qmax, qscale, pmin, pmax, pscale, dev=None,
                 charge=1.0, current=1.0, filling=(1.0,), zoom=1.0,
                 data=None, _handle=None):
        self.dev = dev or Device.default()
        self.filling_set = _f32(filling)
        self.nb = len(self.filling_set)
        self.n = n
        if _handle is not None:
            self.h = _handle
        else:
            h = C.c_void_p()
            d = None if data is None else _f32(data).reshape(self.nb, n, n)
            call("vfps_mesh_create", self.dev.h, self.nb, n, qmin, qmax,
                 qscale, pmin, pmax, pscale, charge, current,
                 _p(self.filling_set), zoom, None if d is None else _p(d),
                 C.byref(h))
            self.h = h
        self.qscale, self.pscale = qscale, pscale

    @property
    def nx(self):
        return self.n

    ny = nx

    def clone(self):
        h = C.c_void_p()
        call("vfps_mesh_clone", self.h, C.byref(h))
        ps = PhaseSpace.__new__(PhaseSpace)
        ps.dev, ps.filling_set, ps.nb, ps.n = (self.dev, self.filling_set,
                                              self.nb, self.n)
        ps.qscale, ps.pscale, ps.h = self.qscale, self.pscale, h
        return ps

    def getData(self):
        out = np.empty((self.nb, self.n, self.n), np.float32)
        call("vfps_mesh_download", self.h, _p(out))
        return out

    def setData(self, data):
        d = _f32(data).reshape(self.nb, self.n, self.n)
        call("vfps_mesh_upload", self.h, _p(d))

    def assign(self, other):
        call("vfps_mesh_copy", self.h, other.h)

    def updateXProjection(self):
        call("vfps_mesh_project_x", self.h)

    def updateYProjection(self):
        call("vfps_mesh_project_y", self.h)

    def getProjection(self, axis):
        out = np.empty((self.nb, self.n), np.float32)
        call("vfps_mesh_get_projection", self.h, axis, _p(out))
        return out

    def setProjection(self, axis, proj):
        p = _f32(proj).reshape(self.nb, self.n)
        call("vfps_mesh_set_projection", self.h, axis, _p(p))

    def integrate(self):
        """Returns (bunch populations, integral), accumulated in FP64."""
        fill = np.empty(self.nb, np.float64)
        integ = C.c_double()
        call("vfps_mesh_integrate", self.h, _p(fill), C.byref(integ))
        return fill, integ.value

    def getIntegral(self):
        return self.integrate()[1]

    def normalize(self):
        call("vfps_mesh_normalize", self.h)

    def integrateAndNormalize(self):
        r = self.integrate()
        self.normalize()
        return r

    def variance(self, axis):
        mean = np.empty(self.nb, np.float32)
        var = np.empty(self.nb, np.float32)
        rms = np.empty(self.nb, np.float32)
        call("vfps_mesh_variance", self.h, axis, _p(mean), _p(var), _p(rms))
        return mean, var, rms

    def _info(self):
        vals = [C.c_uint32(), C.c_uint32(), C.c_float(), C.c_float(),
                C.c_float(), C.c_float()]
        call("vfps_mesh_info", self.h, *[C.byref(v) for v in vals])
        return [v.value for v in vals]

    def getDelta(self, axis):
        return np.float32(self._info()[2 + axis])

    def zerobin(self, axis):
        return np.float32(self._info()[4 + axis])

    def getAxis(self, axis):
        out = np.empty(self.n, np.float32)
        call("vfps_mesh_axis", self.h, axis, _p(out))
        return out

    def simpsonWeights(self):
        out = np.empty(self.n, np.float32)
        call("vfps_mesh_simpson_weights", self.h, _p(out))
        return out

    def device_ptr(self):
        p = C.c_void_p()
        call("vfps_mesh_device_ptr", self.h, C.byref(p))
        return p.value

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib.vfps_mesh_destroy(h)
            self.h = None


class SourceMap:
    """inc/SM/SourceMap.hpp:53-87: apply() runs the map on the device."""

    def __init__(self, h, src, dst):
        self.h = h
        self._keep = (src, dst)

    def apply(self):
        call("vfps_map_apply", self.h)

    def applyToAll(self, positions, seed=0):
        """SourceMap::applyToAll: the map's applyTo on [n][2] grid
        coordinates (x, y); returns the moved positions."""
        p = np.ascontiguousarray(positions, np.float32).reshape(-1, 2).copy()
        call("vfps_map_apply_to", self.h, _p(p), p.shape[0], seed)
        return p

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib.vfps_map_destroy(h)
            self.h = None


class KickMap(SourceMap):
    """src/SM/KickMap.cpp; axis Axis.x (per-y shift along x) or Axis.y."""

    def __init__(self, src, dst, it, clamp, axis):
        h = C.c_void_p()
        call("vfps_map_create_kick", src.h, dst.h, it, int(clamp), axis,
             C.byref(h))
        super().__init__(h, src, dst)
        self.n_off = src.n * src.nb

    def swapOffset(self, offset):
        o = _f32(offset)
        call("vfps_map_set_offset", self.h, _p(o), o.size)

    def getForce(self):
        out = np.empty(self.n_off, np.float32)
        call("vfps_map_get_offset", self.h, _p(out), out.size)
        return out


class DriftMap(KickMap):
    """src/SM/DriftMap.cpp:11-38."""

    def __init__(self, src, dst, slip, E0, it, clamp=False):
        h = C.c_void_p()
        s = _f32(slip)
        call("vfps_map_create_drift", src.h, dst.h, _p(s), s.size, E0, it,
             int(clamp), C.byref(h))
        SourceMap.__init__(self, h, src, dst)
        self.n_off = src.n * src.nb


class RFKickMap(KickMap):
    """src/SM/RFKickMap.cpp: linear (angle, f_RF) or sinusoidal."""

    def __init__(self, src, dst, angle=None, f_RF=None, it=4, clamp=False,
                 revolutionpart=None, V_RF=None, V0=None):
        h = C.c_void_p()
        if revolutionpart is None:
            call("vfps_map_create_rfkick_linear", src.h, dst.h, angle, f_RF,
                 it, int(clamp), C.byref(h))
        else:
            call("vfps_map_create_rfkick", src.h, dst.h, revolutionpart,
                 V_RF, f_RF, V0, it, int(clamp), C.byref(h))
        SourceMap.__init__(self, h, src, dst)
        self.n_off = src.n * src.nb


class DynamicRFKickMap(KickMap):
    """src/SM/DynamicRFKickMap.cpp: an RF kick recomputed before every
    apply() from a precomputed queue of (phase, amplitude) modulations.
    Sinusoidal RF (revolutionpart, V_RF, f_RF, V0) unless `angle` is given
    (linear RF).  seed: the noise generator's seed (the reference draws it
    from std::random_device)."""

    def __init__(self, src, dst, revolutionpart, f_RF, phasespread,
                 amplspread, modampl, modtimeincrement, steps, seed=0,
                 V_RF=None, V0=None, angle=None, it=4, clamp=False):
        h = C.c_void_p()
        if angle is not None:
            call("vfps_map_create_dynrfkick_linear", src.h, dst.h, angle,
                 revolutionpart, f_RF, phasespread, amplspread, modampl,
                 modtimeincrement, steps, seed, it, int(clamp), C.byref(h))
        else:
            call("vfps_map_create_dynrfkick", src.h, dst.h, revolutionpart,
                 V_RF, f_RF, V0, phasespread, amplspread, modampl,
                 modtimeincrement, steps, seed, it, int(clamp), C.byref(h))
        SourceMap.__init__(self, h, src, dst)
        self.n_off = src.n * src.nb
        self.steps = steps
        self._read = 0

    def modulation(self, first=0, count=None):
        """Queue entries [first, first+count) as [count][2] (phase,
        amplitude), and the number applied so far."""
        count = self.steps - first if count is None else count
        out = np.empty((count, 2), np.float32)
        applied = C.c_uint32()
        call("vfps_map_get_modulation", self.h, first, count, _p(out),
             C.byref(applied))
        return out, applied.value

    def getPastModulation(self):
        """Modulations applied since the last call
        (src/SM/DynamicRFKickMap.cpp:113-119)."""
        _, applied = self.modulation(0, 0)
        out, _ = self.modulation(self._read, applied - self._read)
        self._read = applied
        return out


def dynrf_modulation(steps, seed, syncphase, revolutionpart, phasespread,
                     amplspread, modampl, modtimeincrement):
    """The modulation queue of DynamicRFKickMap alone (host only)."""
    out = np.empty((steps, 2), np.float32)
    call("vfps_dynrf_modulation", steps, seed, syncphase, revolutionpart,
         phasespread, amplspread, modampl, modtimeincrement, _p(out))
    return out


class RotationMap(SourceMap):
    """src/SM/RotationMap.cpp:12-176."""

    def __init__(self, src, dst, angle, it, clamp, rotmapsize):
        h = C.c_void_p()
        call("vfps_map_create_rotation", src.h, dst.h, angle, it, int(clamp),
             rotmapsize, C.byref(h))
        super().__init__(h, src, dst)


class FokkerPlanckMap(SourceMap):
    """src/SM/FokkerPlanckMap.cpp:11-212."""

    def __init__(self, src, dst, fptype, fptrack, e1, deriv):
        h = C.c_void_p()
        call("vfps_map_create_fokkerplanck", src.h, dst.h, fptype, fptrack,
             e1, deriv, C.byref(h))
        super().__init__(h, src, dst)


class Identity(SourceMap):
    """inc/SM/Identity.hpp."""

    def __init__(self, src, dst):
        h = C.c_void_p()
        call("vfps_map_create_identity", src.h, dst.h, C.byref(h))
        super().__init__(h, src, dst)


class WakePotentialMap(KickMap):
    """src/SM/WakePotentialMap.cpp: update() then apply()."""

    def __init__(self, src, dst, field, it, clamp=False):
        h = C.c_void_p()
        call("vfps_map_create_wakepotential", src.h, dst.h, field.h, it,
             int(clamp), C.byref(h))
        SourceMap.__init__(self, h, src, dst)
        self.field = field
        self.n_off = src.n * src.nb

    def update(self):
        call("vfps_map_update", self.h)


class Impedance:
    """inc/Z/Impedance.hpp: a complex<float> table on the device."""

    def __init__(self, nfreqs, f_max, dev=None, table=None):
        self.dev = dev or Device.default()
        h = C.c_void_p()
        if table is None:
            call("vfps_impedance_create", self.dev.h, nfreqs, f_max,
                 C.byref(h))
        else:
            t = np.ascontiguousarray(np.asarray(table, np.complex64))
            call("vfps_impedance_create_table", self.dev.h, _p(t), t.size,
                 f_max, C.byref(h))
            nfreqs = t.size
        self.h = h
        self.n = nfreqs
        self.f_max = f_max

    def nFreqs(self):
        return self.n

    def data(self):
        out = np.empty(self.n, np.complex64)
        call("vfps_impedance_download", self.h, _p(out))
        return out

    def __getitem__(self, i):
        return self.data()[i]

    def __iadd__(self, other):
        call("vfps_impedance_add", self.h, other.h)
        return self

    def add_freespace_csr(self, f_rev, f_max):
        call("vfps_impedance_add_freespace_csr", self.h, f_rev, f_max)
        return self

    def add_parallel_plates(self, f0, f_max, g):
        call("vfps_impedance_add_parallel_plates", self.h, f0, f_max, g)
        return self

    def add_resistive_wall(self, f0, f_max, L, s, xi, b):
        call("vfps_impedance_add_resistive_wall", self.h, f0, f_max, L, s,
             xi, b)
        return self

    def add_const(self, z):
        z = complex(z)
        call("vfps_impedance_add_const", self.h, z.real, z.imag)
        return self

    def add_collimator(self, outer, inner):
        call("vfps_impedance_add_collimator", self.h, outer, inner)
        return self

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib.vfps_impedance_destroy(h)
            self.h = None


def FreeSpaceCSR(n, f_rev, f_max, dev=None):  # src/Z/FreeSpaceCSR.cpp
    return Impedance(n, f_max, dev).add_freespace_csr(f_rev, f_max)


def ParallelPlatesCSR(n, f0, f_max, g, dev=None):  # src/Z/ParallelPlatesCSR.cpp
    return Impedance(n, f_max, dev).add_parallel_plates(f0, f_max, g)


def ResistiveWall(n, f0, f_max, L, s, xi, b, dev=None):  # src/Z/ResistiveWall.cpp
    return Impedance(n, f_max, dev).add_resistive_wall(f0, f_max, L, s, xi, b)


def ConstImpedance(n, f_max, z, dev=None):  # src/Z/ConstImpedance.cpp
    return Impedance(n, f_max, dev).add_const(z)


def CollimatorImpedance(n, f_max, outer, inner, dev=None):
    return Impedance(n, f_max, dev).add_collimator(outer, inner)


def makeImpedance(nfreqs, fmax, R_bend, frev, gap, use_csr=True, s=0.0,
                  xi=0.0, inner_coll_radius=0.0, table=None, dev=None):
    """src/Z/ImpedanceFactory.cpp:22-84 (table stands for impedance_file,
    already read).  Returns None when no contribution is selected."""
    f0 = 2.99792458e8 / (2 * np.pi * R_bend)
    rv = Impedance(nfreqs, fmax, dev)
    changed = False
    if gap != 0:
        if use_csr:
            changed = True
            if gap > 0:
                rv.add_parallel_plates(f0, fmax, gap)
            else:
                rv.add_freespace_csr(f0, fmax)
        radius = abs(gap / 2)
        if s > 0 and xi >= -1:
            changed = True
            rv.add_resistive_wall(frev, fmax, 2.99792458e8 / frev, s, xi,
                                  radius)
        if 0 < inner_coll_radius < radius:
            changed = True
            rv.add_collimator(radius, inner_coll_radius)
    if table is not None:
        changed = True
        rv += Impedance(nfreqs, fmax, dev, table=table)
    return rv if changed else None


def wake_potential_noalloc(proj_ptr, nb, n, z_ptr, nmax, buckets,
                           spacing_bins, wakescaling, wake_ptr, dev=None):
    """The wake potential of device profiles (proj_ptr: [nb][n] float32)
    with the device impedance z_ptr (nmax complex64) into wake_ptr ([nb][n]
    float32), allocating nothing (vfps_wake_potential_noalloc).  Pointers
    are device addresses (integers); asynchronous on the device stream."""
    dev = dev or Device.default()
    b = np.ascontiguousarray(buckets, np.uint32)
    call("vfps_wake_potential_noalloc", dev.h, C.c_void_p(proj_ptr), nb, n,
         C.c_void_p(z_ptr), nmax, _p(b), spacing_bins, wakescaling,
         C.c_void_p(wake_ptr))


class ElectricField:
    """inc/PS/ElectricField.hpp."""

    def __init__(self, ps, impedance, buckets, spacing_bins, f_rev=1.0,
                 revolutionpart=1.0, wakescaling=0.0):
        b = np.ascontiguousarray(buckets, np.uint32)
        h = C.c_void_p()
        call("vfps_field_create", ps.h, impedance.h, _p(b), b.size,
             spacing_bins, f_rev, revolutionpart, wakescaling, C.byref(h))
        self.h = h
        self.ps = ps
        self.z = impedance
        self.nmax = impedance.n

    @staticmethod
    def scaling(Ib, dt, qscale, delta1, sigma_delta, E0):
        return np.float32(lib.vfps_field_wakescaling(Ib, dt, qscale, delta1,
                                                     sigma_delta, E0))

    def padBunchProfiles(self):
        call("vfps_field_pad", self.h)

    def getPaddedBunchProfiles(self):
        out = np.empty(self.nmax, np.float32)
        call("vfps_field_get_padded", self.h, _p(out))
        return out

    def getPaddedWakePotential(self):
        out = np.empty(self.nmax, np.float32)
        call("vfps_field_get_padded_wake", self.h, _p(out))
        return out

    def wakePotential(self):
        out = np.empty((self.ps.nb, self.ps.n), np.float32)
        call("vfps_field_wake_potential", self.h, _p(out))
        return out

    def getWakeScaling(self):
        s = C.c_float()
        call("vfps_field_get_wakescaling", self.h, C.byref(s))
        return np.float32(s.value)

    def updateCSR(self, fc):
        power = np.empty(self.ps.nb, np.float32)
        spec = np.empty((self.ps.nb, self.nmax), np.float32)
        call("vfps_field_update_csr", self.h, fc, _p(power), _p(spec))
        return power, spec

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib.vfps_field_destroy(h)
            self.h = None


def airy(x, dev=None):
    """FP64 Ai, Ai', Bi, Bi' on the device (the ParallelPlatesCSR kernel's
    evaluator); returns (values[n,4], overflow flags[n])."""
    dev = dev or Device.default()
    xs = np.ascontiguousarray(x, np.float64).ravel()
    out = np.empty((xs.size, 4), np.float64)
    ovf = np.empty(xs.size, np.int32)
    call("vfps_airy", dev.h, _p(xs), xs.size, _p(out), _p(ovf))
    return out, ovf.astype(bool)


_RING_DEFAULTS = dict(  # src/IO/ProgramOptions.cpp:13-305
    grid_size=256, pq_size=12.0, shift_x=0.0, shift_y=0.0,
    energy_spread=4.7e-4, beam_energy=1.3e9, bending_radius=-1.0,
    f_rev=9e6, harmonic_number=50.0, rf_voltage=1e6, f_s=0.0, alpha0=4e-3,
    alpha1=0.0, alpha2=0.0, vacuum_gap=0.03, damping_time=-1.0, padding=8.0,
    round_padding=1, steps_per_ts=1000, rotations=5.0, bunch_current=3e-3,
    cutoff_frequency=23e9, steps_per_trev=0.0)


def derive(**ring):
    """src/main.cpp:179-355: ring parameters -> derived step parameters."""
    p = dict(_RING_DEFAULTS)
    p.update(ring)
    rp = RingParams(**p)
    d = Derived()
    call("vfps_derive", C.byref(rp), C.byref(d))
    out = {k: getattr(d, k) for k, _ in Derived._fields_}
    out["slip"] = list(d.slip)
    out["ring"] = p
    return out


def toeplitz_kernels(nmax, n, z=None, zc=None, renorm=None, fdelta=0.0):
    """FP64 Toeplitz kernels (g, h) of the direct wake and CSR forms, index
    d + n - 1 for |d| < n (vfps_toeplitz_kernels; host only)."""
    g = np.zeros(2 * n - 1, np.float64) if z is not None else None
    h = np.zeros(2 * n - 1, np.float64) if zc is not None else None
    zz = np.ascontiguousarray(z, np.complex64) if z is not None else None
    zcc = np.ascontiguousarray(zc, np.complex64) if zc is not None else None
    rn = _f32(renorm) if renorm is not None else None
    call("vfps_toeplitz_kernels", nmax, n,
         _p(zz) if zz is not None else None,
         _p(zcc) if zcc is not None else None,
         _p(rn) if rn is not None else None, float(fdelta),
         _p(g) if g is not None else None, _p(h) if h is not None else None)
    return g, h


def _step_desc(desc, out=None):
    d = out if out is not None else StepDesc()
    d.interp, d.fused, d.nslip, d.use_fp, d.fptype = 4, 1, 3, 1, 3
    d.deriv = 4
    for k, v in desc.items():
        if k == "slip":
            s = list(v) + [0.0] * (3 - len(v))
            d.slip = (C.c_float * 3)(*s)
            d.nslip = len(v)
        else:
            setattr(d, k, v)
    return d


class Step:
    """The time loop of src/main.cpp:948-1053 on one mesh (see
    vfps_step_desc in include/vfps_gpu.h for the fields of desc)."""

    def __init__(self, desc, init, zwake=None, zcsr=None, dev=None):
        self.dev = dev or Device.default()
        d = _step_desc(desc)
        self.desc = d
        self.n = d.n
        init = _f32(init).reshape(self.n, self.n)
        h = C.c_void_p()
        call("vfps_step_create", self.dev.h, C.byref(d), _p(init),
             zwake.h if zwake else None, zcsr.h if zcsr else None,
             C.byref(h))
        self.h = h
        self._keep = (zwake, zcsr)

    def run(self, nsteps):
        call("vfps_step_run", self.h, nsteps)

    def prepare(self, nsteps):
        """Capture the graphs a run of nsteps from the current state replays
        (a timed run then spends no time capturing)."""
        call("vfps_step_prepare", self.h, nsteps)

    def record(self, capacity, profiles=False):
        call("vfps_step_record", self.h, capacity, int(profiles))
        self.cap = capacity

    def history(self):
        cnt = C.c_uint32()
        cap = getattr(self, "cap", 0)
        csr = np.zeros(cap, np.float32)
        prof = np.zeros((cap, self.n), np.float32)
        call("vfps_step_get_history", self.h, C.byref(cnt), _p(csr),
             _p(prof))
        c = cnt.value
        return csr[:c], prof[:c]

    def mesh(self):
        m = np.empty((self.n, self.n), np.float32)
        p = np.empty(self.n, np.float32)
        w = np.empty(self.n, np.float32)
        call("vfps_step_get_mesh", self.h, _p(m), _p(p), _p(w))
        return m, p, w

    def set_particles(self, positions, fptrack=3, seed=0):
        """Track [n][2] grid coordinates through every following step."""
        p = np.ascontiguousarray(positions, np.float32).reshape(-1, 2)
        call("vfps_step_set_particles", self.h, _p(p), p.shape[0], fptrack,
             seed)
        self.npart = p.shape[0]

    def particles(self):
        out = np.empty((getattr(self, "npart", 0), 2), np.float32)
        call("vfps_step_get_particles", self.h, _p(out), out.shape[0])
        return out

    def moments(self):
        """Outstep quantities of the current state (src/main.cpp:961-967):
        dict of xproj, yproj (n each) and population, position, length,
        average, spread (floats)."""
        n = self.n
        xp = np.empty(n, np.float32)
        yp = np.empty(n, np.float32)
        v = [C.c_float() for _ in range(5)]
        call("vfps_step_moments", self.h, _p(xp), _p(yp),
             *[C.byref(x) for x in v])
        names = ("population", "position", "length", "average", "spread")
        out = {k: np.float32(x.value) for k, x in zip(names, v)}
        out.update(xproj=xp, yproj=yp)
        return out

    def set_mesh(self, mesh):
        m = _f32(mesh).reshape(self.n, self.n)
        call("vfps_step_set_mesh", self.h, _p(m))

    def timing(self, nsteps):
        a, b, c = C.c_float(), C.c_float(), C.c_float()
        call("vfps_step_timing", self.h, nsteps, C.byref(a), C.byref(b),
             C.byref(c))
        return a.value, b.value, c.value

    def phases(self, tiles=False):
        """Device timestamps of one fused step (ns): the step kernel's phase
        ends in its first tile, the field kernel's phase ends (FFT wake
        only), and the step workgroups' (gap after the field kernel, entry
        spread, shortest, median, longest, last exit); with tiles=True also
        every tile's (entry, exit) after the first entry."""
        ntiles = -(-self.n // self.config()["tile_rows"])
        out = np.zeros(24 + 2 * ntiles, np.float64)
        call("vfps_step_phases", self.h, _p(out), out.size)
        r = (out[1:8].copy(), out[9:14].copy(), out[16:22].copy())
        if tiles:
            r += (out[24:].reshape(ntiles, 2).copy(),)
        return r

    def config(self):
        """The configuration in use: fused, wake method (1 FFT, 2 direct),
        tile rows, threads, LDS buffer rows, LDS bytes."""
        out = np.zeros(6, np.uint32)
        call("vfps_step_config", self.h, _p(out), 6)
        return dict(zip(("fused", "wake_method", "tile_rows", "threads",
                         "stride", "lds_bytes"), (int(v) for v in out)))

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib.vfps_step_destroy(h)
            self.h = None


class Batch:
    """Independent meshes of any sizes advanced together, one launch per
    step (vfps_batch_* in include/vfps_gpu.h).  descs, inits, zwake, zcsr:
    one entry per instance."""

    def __init__(self, descs, inits, zwake, zcsr, dev=None):
        self.dev = dev or Device.default()
        k = len(descs)
        arr = (StepDesc * k)()
        for i, dsc in enumerate(descs):
            _step_desc(dsc, arr[i])
        self.n = [int(dsc["n"]) for dsc in descs]
        self._inits = [_f32(m).reshape(self.n[i], self.n[i])
                       for i, m in enumerate(inits)]
        ptrs = (C.c_void_p * k)(*[m.ctypes.data for m in self._inits])
        zw = (C.c_void_p * k)(*[z.h if z else None for z in zwake])
        zc = (C.c_void_p * k)(*[z.h if z else None for z in zcsr])
        h = C.c_void_p()
        call("vfps_batch_create", self.dev.h, k, arr, ptrs, zw, zc,
             C.byref(h))
        self.h = h
        self._keep = (list(zwake), list(zcsr))
        self.cap = 0

    def __len__(self):
        return len(self.n)

    def run(self, nsteps):
        call("vfps_batch_run", self.h, nsteps)

    def prepare(self, nsteps):
        """Capture the graphs a run of nsteps from the current state replays."""
        call("vfps_batch_prepare", self.h, nsteps)

    def record(self, capacity, profiles=False):
        call("vfps_batch_record", self.h, capacity, int(profiles))
        self.cap = capacity
        self.profiles = profiles

    def history(self, i):
        cnt = C.c_uint32()
        csr = np.zeros(self.cap, np.float32)
        prof = np.zeros((self.cap, self.n[i]), np.float32) if getattr(
            self, "profiles", False) else None
        call("vfps_batch_get_history", self.h, i, C.byref(cnt), _p(csr),
             _p(prof) if prof is not None else None)
        c = cnt.value
        return csr[:c], (prof[:c] if prof is not None else None)

    def mesh(self, i):
        n = self.n[i]
        m = np.empty((n, n), np.float32)
        p = np.empty(n, np.float32)
        w = np.empty(n, np.float32)
        call("vfps_batch_get_mesh", self.h, i, _p(m), _p(p), _p(w))
        return m, p, w

    def projections(self):
        """Current x projection of every instance (list of arrays)."""
        out = np.empty(sum(self.n), np.float32)
        call("vfps_batch_get_projections", self.h, _p(out))
        return np.split(out, np.cumsum(self.n)[:-1])

    def csr_histories(self, first, count):
        """Recorded CSR power of steps [first, first+count), [ninst][count]."""
        out = np.empty((len(self.n), count), np.float32)
        call("vfps_batch_get_csr_histories", self.h, first, count, _p(out))
        return out

    def info(self):
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint64()
        call("vfps_batch_info", self.h, C.byref(a), C.byref(b), C.byref(c))
        return dict(instances=a.value, launch_groups=b.value,
                    cells=c.value)

    def config(self, i):
        """Instance i's configuration: fused, wake method, tile rows,
        threads, LDS buffer rows, LDS bytes, persistent launch, tiles per
        strip (0 outside the strip form)."""
        out = np.zeros(8, np.uint32)
        call("vfps_batch_config", self.h, i, _p(out), 8)
        return dict(zip(("fused", "wake_method", "tile_rows", "threads",
                         "stride", "lds_bytes", "persistent", "strip"),
                        (int(v) for v in out)))

    def launches(self):
        """The launches of one step: threads, tiles, workgroups, LDS bytes
        (fewer workgroups than tiles: a persistent launch)."""
        cap = 64
        out = np.zeros((cap, 4), np.uint32)
        ng = C.c_uint32()
        call("vfps_batch_launches", self.h, _p(out), cap, C.byref(ng))
        return [dict(threads=int(r[0]), tiles=int(r[1]), workgroups=int(r[2]),
                     lds_bytes=int(r[3])) for r in out[:min(ng.value, cap)]]

    def set_merge(self, merge):
        """One launch per step for all size classes (merge=True) or one per
        LDS class as parallel branches (False, the default)."""
        call("vfps_batch_set_merge", self.h, 1 if merge else 0)

    def retile_size(self, n, tile_rows, threads, persistent=False, strip=0):
        """Set the tiling of every instance of mesh size n (strip > 0: the
        strip form, strip tiles per workgroup)."""
        if strip:
            call("vfps_batch_retile_size_strip", self.h, n, tile_rows,
                 threads, strip)
        else:
            call("vfps_batch_retile_size", self.h, n, tile_rows, threads,
                 1 if persistent else 0)

    def retile(self, i, tile_rows, threads, persistent=False, strip=0):
        """Set instance i's tiling by hand (results do not depend on it)."""
        if strip:
            call("vfps_batch_retile_strip", self.h, i, tile_rows, threads,
                 strip)
        else:
            call("vfps_batch_retile", self.h, i, tile_rows, threads,
                 1 if persistent else 0)

    def timing(self, nsteps):
        ms = C.c_float()
        call("vfps_batch_timing", self.h, nsteps, C.byref(ms))
        return ms.value

    def autotune(self, steps=20):
        """Pick the tile height, workgroup size and launch form (one
        workgroup per tile, persistent, or strips of tiles) per mesh size;
        returns the timing table (n, T, threads, us/step, chosen,
        persistent, retimed: the three fastest of a size are timed again
        over 4 x steps; strip: tiles per strip, 0 outside the strip
        form)."""
        cap = 1024
        table = np.zeros((cap, 8), np.float32)
        rows = C.c_uint32()
        call("vfps_batch_autotune", self.h, steps, _p(table), cap,
             C.byref(rows))
        return [dict(n=int(r[0]), T=int(r[1]), threads=int(r[2]),
                     us_per_step=float(r[3]), chosen=bool(r[4]),
                     persistent=bool(r[5]), retimed=bool(r[6]),
                     strip=int(r[7]))
                for r in table[:min(rows.value, cap)]]

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib.vfps_batch_destroy(h)
            self.h = None


def round_bf16(x):
    """The kernels' BF16 rounding (nearest even), as float32."""
    x = _f32(x)
    y = np.empty_like(x)
    call("vfps_round_bf16", _p(x), _p(y), x.size)
    return y


def gemm_bf16(A, B, dev=None):
    """A @ B on the matrix cores (operands rounded to BF16, FP32
    accumulation); A, B 2-D arrays (any strides are taken from numpy)."""
    dev = dev or Device.default()
    A = np.asarray(A, np.float32)
    B = np.asarray(B, np.float32)
    M, K = A.shape
    K2, N = B.shape
    assert K == K2
    a = np.ascontiguousarray(A)
    b = np.ascontiguousarray(B)
    C_ = np.empty((M, N), np.float32)
    call("vfps_gemm_bf16", dev.h, M, N, K, _p(a), a.size, K, 1, _p(b),
         b.size, N, 1, _p(C_))
    return C_


class MLP:
    """The surrogate network (vfps_mlp_* in include/vfps_gpu.h): dims =
    (n_in, hidden..., n_out); GELU after every hidden layer; the last output
    is a logit."""

    def __init__(self, dims, max_batch=256, seed=0, dev=None):
        self.dev = dev or Device.default()
        self.dims = tuple(int(d) for d in dims)
        arr = np.asarray(self.dims, np.uint32)
        h = C.c_void_p()
        call("vfps_mlp_create", self.dev.h, _p(arr), len(self.dims),
             max_batch, seed, C.byref(h))
        self.h = h
        n = C.c_size_t()
        call("vfps_mlp_nparams", self.h, C.byref(n))
        self.nparams = n.value
        self.max_batch = max_batch

    def layers(self, flat):
        """Split a flat parameter vector into [(W, b), ...]."""
        out, o = [], 0
        for a, b in zip(self.dims[:-1], self.dims[1:]):
            W = flat[o:o + a * b].reshape(b, a)
            o += a * b
            out.append((W, flat[o:o + b]))
            o += b
        return out

    def params(self):
        p = np.empty(self.nparams, np.float32)
        call("vfps_mlp_get_params", self.h, _p(p), None, self.nparams)
        return p

    def grads(self):
        g = np.empty(self.nparams, np.float32)
        call("vfps_mlp_get_params", self.h, None, _p(g), self.nparams)
        return g

    def set_params(self, p):
        p = _f32(p)
        call("vfps_mlp_set_params", self.h, _p(p), p.size)

    def forward(self, x):
        x = _f32(x).reshape(-1, self.dims[0])
        y = np.empty((x.shape[0], self.dims[-1]), np.float32)
        call("vfps_mlp_forward", self.h, _p(x), x.shape[0], _p(y))
        return y

    def loss_grad(self, x, target, label, label_weight=1.0):
        x = _f32(x).reshape(-1, self.dims[0])
        t = _f32(target).reshape(x.shape[0], self.dims[-1] - 1)
        lab = _f32(label).reshape(x.shape[0])
        loss = C.c_float()
        call("vfps_mlp_loss_grad", self.h, _p(x), _p(t), _p(lab), x.shape[0],
             label_weight, C.byref(loss))
        return loss.value

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        call("vfps_mlp_adam_step", self.h, lr, beta1, beta2, eps)

    def train(self, x, target, label, batch=256, epochs=10, lr=1e-3,
              label_weight=1.0, seed=0):
        x = _f32(x).reshape(-1, self.dims[0])
        t = _f32(target).reshape(x.shape[0], self.dims[-1] - 1)
        lab = _f32(label).reshape(x.shape[0])
        losses = np.zeros(epochs, np.float32)
        call("vfps_mlp_train", self.h, _p(x), _p(t), _p(lab), x.shape[0],
             batch, epochs, lr, label_weight, seed, _p(losses))
        return losses

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib.vfps_mlp_destroy(h)
            self.h = None
