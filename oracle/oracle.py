"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg to
check and time the HIP library.  The product package never imports this.

``Oracle(fp64=False)`` wraps liboracle_f32.so (the restatement of the
reference CPU path, parity target); ``Oracle(fp64=True)`` wraps
liboracle_f64.so (same code with every float type widened to double).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_u32 = C.c_uint32
_i32 = C.c_int32
_f64 = C.c_double
_vp = C.c_void_p


def _ptr(a):
    return a.ctypes.data_as(_vp) if a is not None else None


class SimCfg(C.Structure):
    pass


_DYNRF = None


def _dynrf_lib():
    global _DYNRF
    if _DYNRF is None:
        path = os.path.join(HERE, "liboracle_dynrf.so")
        if not os.path.exists(path):
            from oracle.build_oracle import build
            build()
        L = C.CDLL(path)
        L.or_dynrf_modulation.argtypes = [_u32, _u32, C.c_float, _f64,
                                          C.c_float, C.c_float, C.c_float,
                                          _f64, _vp]
        L.or_mt19937_10000th.restype = _u32
        _DYNRF = L
    return _DYNRF


class Oracle:
    def __init__(self, fp64=False, path=None):
        """path: another build of the same source (the natively compiled
        CPU baseline of the benchmark)."""
        name = "liboracle_f64.so" if fp64 else "liboracle_f32.so"
        path = path or os.path.join(HERE, name)
        if not os.path.exists(path):
            from oracle.build_oracle import build
            build()
        self.lib = C.CDLL(path)
        self.fp64 = fp64
        self.real = np.float64 if fp64 else np.float32
        self.creal = C.c_double if fp64 else C.c_float
        self.hi = np.dtype([("index", np.uint32), ("weight", self.real)],
                           align=True)
        assert self.hi.itemsize == self.lib.or_sizeof_hi()
        self.cplx = np.dtype([("re", self.real), ("im", self.real)])
        R = self.creal

        class Cfg(C.Structure):
            _fields_ = [("n", _u32), ("qmin", R), ("qmax", R), ("pmin", R),
                        ("pmax", R), ("qscale", R), ("pscale", R),
                        ("scheme", _u32), ("interp", _u32),
                        ("rot_clamp", _i32), ("angle", R),
                        ("rf_linear", _i32), ("revolutionpart", R),
                        ("V_RF", R), ("f_RF", R), ("V0", R),
                        ("slip", R * 3), ("nslip", _u32), ("E0", R),
                        ("use_fp", _i32), ("fptype", _u32), ("e1", R),
                        ("deriv", _u32), ("use_wake", _i32), ("nmax", _u32),
                        ("spacing_bins", _u32), ("wakescaling_arg", R),
                        ("use_csr", _i32), ("fc", R),
                        ("renormalize", _u32)]
        self.Cfg = Cfg
        L = self.lib
        L.or_sim_create.restype = _vp
        L.or_sim_create.argtypes = [C.POINTER(Cfg), _vp, _vp, _vp]
        L.or_sim_destroy.argtypes = [_vp]
        L.or_sim_run.argtypes = [_vp, _u32, _vp, _vp]
        L.or_sim_get.argtypes = [_vp, _vp, _vp, _vp]
        L.or_sim_set_mesh.argtypes = [_vp, _vp]
        L.or_ruler.argtypes = [_u32, R, R, _vp, _vp, _vp]
        L.or_upper_power_of_two.restype = C.c_uint64
        L.or_upper_power_of_two.argtypes = [C.c_uint64]
        L.or_simpson_weights.argtypes = [_u32, R, _vp]
        L.or_gaus.argtypes = [_u32, _vp, _f64, _vp]
        L.or_phasespace_gaussian.restype = C.c_int
        L.or_phasespace_gaussian.argtypes = [_u32, _u32, R, R, R, R, _vp,
                                             _f64, _vp, _vp, _vp, _vp, _vp]
        L.or_filling_ok.restype = C.c_int
        L.or_filling_ok.argtypes = [_u32, _vp]
        for f in ("or_project_x", "or_project_y"):
            getattr(L, f).argtypes = [_u32, _u32, _vp, _vp, _vp]
        L.or_integrate.argtypes = [_u32, _u32, _vp, _vp, _vp, _vp]
        L.or_integrate_f64.argtypes = [_u32, _u32, _vp, _vp, _vp, _vp]
        L.or_normalize.argtypes = [_u32, _u32, _vp, _vp, _vp]
        L.or_variance.argtypes = [_u32, _u32, _vp, _vp, R, _vp, _vp, _vp,
                                  _vp, _vp]
        L.or_calc_coefficients.argtypes = [_vp, R, _u32]
        L.or_sourcemap_apply.argtypes = [_u32, _u32, _vp, _vp, _vp]
        L.or_kick_update_sm.argtypes = [_u32, _u32, _vp, _u32, _vp]
        L.or_kick_apply_x.argtypes = [_u32, _u32, _u32, _u32, _vp, _vp, _vp]
        L.or_kick_apply_y.argtypes = [_u32, _u32, _u32, _u32, _vp, _vp, _vp]
        L.or_drift_offsets.argtypes = [_u32, _vp, R, _vp, _u32, R, R, _vp]
        L.or_rfkick_offsets_linear.argtypes = [_u32, R, R, R, R, R, _vp]
        L.or_rfkick_offsets.argtypes = [_u32, _vp, R, R, R, R, R, R, R, _vp]
        L.or_rfkick_offsets_at.argtypes = [_u32, _vp, R, R, R, R, R, R, R,
                                           R, R, _vp]
        L.or_track_kick.argtypes = [C.c_int, _u32, _u32, _vp, _vp]
        L.or_track_rotation.argtypes = [R, R, R, R, _vp]
        L.or_track_fp.argtypes = [C.c_int, _u32, _u32, _u32, _vp, _vp, _vp]
        L.or_rfkick_offsets_linear_at.argtypes = [_u32, R, R, R, R, R, R, R,
                                                  R, _vp]
        L.or_rotation_hinfo.argtypes = [_u32, _u32, _vp, R, R, _vp, R, R, R,
                                        R, _u32, _u32, _u32, _vp]
        L.or_rotation_build.restype = C.c_int
        L.or_rotation_build.argtypes = [_u32, _vp, R, R, _vp, R, R, R, _u32,
                                        C.c_int, _u32, _vp, _vp, _vp]
        L.or_rotation_apply.argtypes = [_u32, _u32, C.c_int, _u32, _vp, _vp,
                                        R, R, _vp, R, R, R, R, _vp, _vp, _vp]
        L.or_fp_build.restype = C.c_int
        L.or_fp_build.argtypes = [_u32, _u32, R, _u32, _vp, R, R, _vp]
        L.or_fp_apply.argtypes = [_u32, _u32, _u32, _vp, _vp, _vp]
        L.or_fft_r2c.argtypes = [_u32, _vp, _vp]
        L.or_fft_c2r.argtypes = [_u32, _vp, _vp]
        L.or_dft_r2c_f64.argtypes = [_u32, _vp, _vp, _vp]
        L.or_pad_profiles.argtypes = [_u32, _u32, _vp, _u32, _vp, _u32, _vp]
        L.or_wake_potential.argtypes = [_u32, _u32, _vp, _vp, _u32, _vp,
                                        _u32, R, _vp, _vp]
        L.or_update_csr.argtypes = [_u32, _u32, _vp, _vp, _u32, R, R, R, _vp,
                                    _vp]
        L.or_wakescaling.restype = R
        L.or_wakescaling.argtypes = [_f64, _f64, _f64, R, _f64, _f64]
        L.or_z_const.argtypes = [_u32, R, R, _vp]
        L.or_z_freespace.argtypes = [_u32, R, R, _vp]
        L.or_z_resistive_wall.argtypes = [_u32, R, R, _f64, _f64, _f64, _f64,
                                          _vp]
        L.or_z_collimator.argtypes = [_u32, _f64, _f64, _vp]
        L.or_z_parallel_plates.restype = C.c_long
        L.or_z_parallel_plates.argtypes = [_u32, R, R, _f64, _vp]
        L.or_airy.restype = C.c_int
        L.or_airy.argtypes = [_f64, _vp, _vp, _vp, _vp]

    # ---- helpers ------------------------------------------------------
    def arr(self, a):
        return np.ascontiguousarray(a, dtype=self.real)

    def ruler(self, steps, vmin, vmax):
        data = np.empty(steps, self.real)
        d = self.creal()
        z = self.creal()
        self.lib.or_ruler(steps, vmin, vmax, _ptr(data), C.byref(d),
                          C.byref(z))
        return data, self.real(d.value), self.real(z.value)

    def upper_power_of_two(self, v):
        return int(self.lib.or_upper_power_of_two(v))

    def simpson_weights(self, n, delta0):
        ws = np.empty(n, self.real)
        self.lib.or_simpson_weights(n, delta0, _ptr(ws))
        return ws

    def gaus(self, axis, zoom=1.0):
        axis = self.arr(axis)
        out = np.empty_like(axis)
        self.lib.or_gaus(len(axis), _ptr(axis), zoom, _ptr(out))
        return out

    def phasespace_gaussian(self, nb, n, qmin, qmax, pmin, pmax, filling,
                            zoom=1.0):
        fs = self.arr(filling)
        data = np.empty((nb, n, n), self.real)
        px = np.empty((nb, n), self.real)
        py = np.empty((nb, n), self.real)
        fill = np.empty(nb, self.real)
        integ = np.empty(1, self.real)
        rc = self.lib.or_phasespace_gaussian(nb, n, qmin, qmax, pmin, pmax,
                                             _ptr(fs), zoom, _ptr(data),
                                             _ptr(px), _ptr(py), _ptr(fill),
                                             _ptr(integ))
        if rc != 0:
            raise ValueError('Argument "filling" not normalized.')
        return data, px, py, fill, integ[0]

    def project_x(self, data, ws):
        data = self.arr(data)
        nb, n = data.shape[0], data.shape[1]
        out = np.empty((nb, n), self.real)
        self.lib.or_project_x(nb, n, _ptr(data), _ptr(self.arr(ws)),
                              _ptr(out))
        return out

    def project_y(self, data, ws):
        data = self.arr(data)
        nb, n = data.shape[0], data.shape[1]
        out = np.empty((nb, n), self.real)
        self.lib.or_project_y(nb, n, _ptr(data), _ptr(self.arr(ws)),
                              _ptr(out))
        return out

    def integrate(self, proj, ws):
        proj = self.arr(proj)
        nb, n = proj.shape
        fill = np.empty(nb, self.real)
        integ = np.empty(1, self.real)
        self.lib.or_integrate(nb, n, _ptr(proj), _ptr(self.arr(ws)),
                              _ptr(fill), _ptr(integ))
        return fill, integ[0]

    def integrate_f64(self, proj, ws):
        proj = self.arr(proj)
        nb, n = proj.shape
        fill = np.empty(nb, np.float64)
        integ = np.empty(1, np.float64)
        self.lib.or_integrate_f64(nb, n, _ptr(proj), _ptr(self.arr(ws)),
                                  _ptr(fill), _ptr(integ))
        return fill, integ[0]

    def normalize(self, data, filling_set, filling):
        data = self.arr(data).copy()
        nb, n = data.shape[0], data.shape[1]
        self.lib.or_normalize(nb, n, _ptr(data), _ptr(self.arr(filling_set)),
                              _ptr(self.arr(filling)))
        return data

    def variance(self, proj_axis, axis, delta, filling_set, filling):
        proj_axis = self.arr(proj_axis)
        nb, n = proj_axis.shape
        mean = np.empty(nb, self.real)
        var = np.empty(nb, self.real)
        rms = np.empty(nb, self.real)
        self.lib.or_variance(nb, n, _ptr(proj_axis), _ptr(self.arr(axis)),
                             delta, _ptr(self.arr(filling_set)),
                             _ptr(self.arr(filling)), _ptr(mean), _ptr(var),
                             _ptr(rms))
        return mean, var, rms

    def coefficients(self, f, it):
        ic = np.zeros(4, self.real)
        self.lib.or_calc_coefficients(_ptr(ic), f, it)
        return ic[:it]

    def kick_update_sm(self, kd, offset, it):
        offset = self.arr(offset)
        h = np.zeros(len(offset) * it, self.hi)
        self.lib.or_kick_update_sm(kd, len(offset), _ptr(offset), it, _ptr(h))
        return h

    def kick_apply(self, axis, data, hinfo, it):
        """axis 0: x kick (src/SM/KickMap.cpp:221-243), 1: y kick."""
        data = self.arr(data)
        nb, n = data.shape[0], data.shape[1]
        out = np.empty_like(data)
        f = self.lib.or_kick_apply_x if axis == 0 else self.lib.or_kick_apply_y
        f(nb, n, n, it, _ptr(hinfo), _ptr(data), _ptr(out))
        return out

    def kick(self, axis, data, offset, it):
        n = data.shape[-1]
        h = self.kick_update_sm(n, offset, it)
        return self.kick_apply(axis, data, h, it)

    def drift_offsets(self, axis1, delta0, slip, pscale, E0):
        axis1 = self.arr(axis1)
        slip = self.arr(slip)
        out = np.empty_like(axis1)
        self.lib.or_drift_offsets(len(axis1), _ptr(axis1), delta0, _ptr(slip),
                                  len(slip), pscale, E0, _ptr(out))
        return out

    def rfkick_offsets_linear(self, n, zerobin0, angle, qscale, f_RF, delta0):
        out = np.empty(n, self.real)
        self.lib.or_rfkick_offsets_linear(n, zerobin0, angle, qscale, f_RF,
                                          delta0, _ptr(out))
        return out

    def rfkick_offsets(self, axis0, revolutionpart, V_RF, f_RF, V0, qscale,
                       delta1, pscale):
        axis0 = self.arr(axis0)
        out = np.empty_like(axis0)
        self.lib.or_rfkick_offsets(len(axis0), _ptr(axis0), revolutionpart,
                                   V_RF, f_RF, V0, qscale, delta1, pscale,
                                   _ptr(out))
        return out

    def rfkick_offsets_at(self, axis0, revolutionpart, V_RF, f_RF, V0,
                          qscale, delta1, pscale, phase, ampl):
        """RFKickMap::_calcKick(phase, ampl), sinusoidal."""
        axis0 = self.arr(axis0)
        out = np.empty_like(axis0)
        self.lib.or_rfkick_offsets_at(len(axis0), _ptr(axis0), revolutionpart,
                                      V_RF, f_RF, V0, qscale, delta1, pscale,
                                      phase, ampl, _ptr(out))
        return out

    def rfkick_offsets_linear_at(self, n, zerobin0, angle, qscale, f_RF,
                                 delta0, syncphase, phase, ampl):
        """RFKickMap::_calcKick(phase, ampl), linear."""
        out = np.empty(n, self.real)
        self.lib.or_rfkick_offsets_linear_at(n, zerobin0, angle, qscale, f_RF,
                                             delta0, syncphase, phase, ampl,
                                             _ptr(out))
        return out

    def track_kick(self, axis, n, offset, pos):
        """KickMap::applyTo on each [x, y] of pos (src/SM/KickMap.cpp:280-304)."""
        off = self.arr(offset)
        out = np.array(pos, self.real).reshape(-1, 2).copy()
        for p in out:
            self.lib.or_track_kick(axis, n, n, _ptr(off), _ptr(p))
        return out

    def track_rotation(self, cos_dt, sin_dt, zb0, zb1, pos):
        out = np.array(pos, self.real).reshape(-1, 2).copy()
        for p in out:
            self.lib.or_track_rotation(cos_dt, sin_dt, zb0, zb1, _ptr(p))
        return out

    def track_fp(self, mode, hinfo, ip, n, data_in, pos):
        """FokkerPlanckMap::applyTo approximation1/2 (mode 1/2)."""
        out = np.array(pos, self.real).reshape(-1, 2).copy()
        d = self.arr(data_in) if data_in is not None else None
        for p in out:
            self.lib.or_track_fp(mode, ip, n, n, _ptr(hinfo), _ptr(d),
                                 _ptr(p))
        return out

    @staticmethod
    def dynrf_modulation(steps, seed, syncphase, revolutionpart, phasespread,
                         amplspread, modampl, modtimeincrement):
        """DynamicRFKickMap's modulation queue [steps][2] (phase, ampl)."""
        L = _dynrf_lib()
        out = np.empty((steps, 2), np.float32)
        L.or_dynrf_modulation(steps, seed, syncphase, revolutionpart,
                              phasespread, amplspread, modampl,
                              modtimeincrement, _ptr(out))
        return out

    @staticmethod
    def mt19937_10000th():
        return _dynrf_lib().or_mt19937_10000th()

    def rotation(self, data, axis0, d0, zb0, axis1, d1, zb1, angle, it,
                 clamp, rotmapsize):
        """Constructor + apply of RotationMap; raises ValueError where the
        reference constructor throws std::invalid_argument."""
        data = self.arr(data)
        n = data.shape[-1]
        a0 = self.arr(axis0)
        a1 = self.arr(axis1)
        h = np.zeros(n * n * it * it if rotmapsize else 1, self.hi)
        cs = self.creal()
        sn = self.creal()
        rc = self.lib.or_rotation_build(n, _ptr(a0), d0, zb0, _ptr(a1), d1,
                                        zb1, angle, it, int(clamp),
                                        rotmapsize, _ptr(h), C.byref(cs),
                                        C.byref(sn))
        if rc != 0:
            raise ValueError("Clamping only supported with cubic "
                             "interpolation and rotmapsize > 0.")
        scratch = np.zeros(16, self.hi)
        out = np.empty_like(data)
        self.lib.or_rotation_apply(n, it, int(clamp), rotmapsize, _ptr(h),
                                   _ptr(a0), d0, zb0, _ptr(a1), d1, zb1,
                                   cs.value, sn.value, _ptr(scratch),
                                   _ptr(data), _ptr(out))
        return out

    def fp_build(self, n, fptype, e1, deriv, axis1, d1, zb1):
        h = np.zeros(n * 5, self.hi)
        ip = self.lib.or_fp_build(n, fptype, e1, deriv, _ptr(self.arr(axis1)),
                                  d1, zb1, _ptr(h))
        if ip < 0:
            raise ValueError("unknown derivation type")
        return h[: n * ip], ip

    def fp_apply(self, data, hinfo, ip):
        data = self.arr(data)
        nb, n = data.shape[0], data.shape[1]
        out = np.empty_like(data)
        self.lib.or_fp_apply(nb, n, ip, _ptr(hinfo), _ptr(data), _ptr(out))
        return out

    def fft_r2c(self, x):
        x = self.arr(x)
        n = len(x)
        out = np.zeros(n // 2 + 1, self.cplx)
        self.lib.or_fft_r2c(n, _ptr(x), _ptr(out))
        return out["re"].astype(np.float64) + 1j * out["im"]

    def fft_c2r(self, w, n):
        w = np.asarray(w)
        buf = np.zeros(n // 2 + 1, self.cplx)
        buf["re"] = w.real
        buf["im"] = w.imag
        out = np.empty(n, self.real)
        self.lib.or_fft_c2r(n, _ptr(buf), _ptr(out))
        return out

    def dft_r2c_f64(self, x):
        x = np.ascontiguousarray(x, np.float64)
        n = len(x)
        re = np.empty(n // 2 + 1)
        im = np.empty(n // 2 + 1)
        self.lib.or_dft_r2c_f64(n, _ptr(x), _ptr(re), _ptr(im))
        return re + 1j * im

    def to_cplx(self, z):
        z = np.asarray(z)
        buf = np.zeros(len(z), self.cplx)
        buf["re"] = z.real
        buf["im"] = z.imag
        return buf

    @staticmethod
    def from_cplx(buf):
        return buf["re"].astype(np.float64) + 1j * buf["im"].astype(np.float64)

    def pad_profiles(self, proj, nmax, buckets, spacing):
        proj = self.arr(proj)
        nb, n = proj.shape
        b = np.ascontiguousarray(buckets, np.uint32)
        out = np.empty(nmax, self.real)
        self.lib.or_pad_profiles(nb, n, _ptr(proj), nmax, _ptr(b), spacing,
                                 _ptr(out))
        return out

    def wake_potential(self, proj, Z, buckets, spacing, wakescaling_arg):
        proj = self.arr(proj)
        nb, n = proj.shape
        zb = self.to_cplx(Z)
        b = np.ascontiguousarray(buckets, np.uint32)
        wp = np.empty((nb, n), self.real)
        pad = np.empty(len(zb), self.real)
        self.lib.or_wake_potential(nb, n, _ptr(proj), _ptr(zb), len(zb),
                                   _ptr(b), spacing, wakescaling_arg,
                                   _ptr(wp), _ptr(pad))
        return wp, pad

    def update_csr(self, proj, Z, delta0, qscale, fc):
        proj = self.arr(proj)
        nb, n = proj.shape
        zb = self.to_cplx(Z)
        nmax = len(zb)
        spec = np.empty((nb, nmax), self.real)
        power = np.empty(nb, self.real)
        self.lib.or_update_csr(nb, n, _ptr(proj), _ptr(zb), nmax, delta0,
                               qscale, fc, _ptr(spec), _ptr(power))
        return spec, power

    def wakescaling(self, Ib, dt, qscale, delta1, sigma_delta, E0):
        return self.real(self.lib.or_wakescaling(Ib, dt, qscale, delta1,
                                                 sigma_delta, E0))

    def z_const(self, n, z):
        out = np.zeros(n, self.cplx)
        self.lib.or_z_const(n, z.real, z.imag, _ptr(out))
        return self.from_cplx(out)

    def z_freespace(self, n, f_rev, f_max):
        out = np.zeros(n, self.cplx)
        self.lib.or_z_freespace(n, f_rev, f_max, _ptr(out))
        return self.from_cplx(out)

    def z_resistive_wall(self, n, f0, f_max, L, s, xi, b):
        out = np.zeros(n, self.cplx)
        self.lib.or_z_resistive_wall(n, f0, f_max, L, s, xi, b, _ptr(out))
        return self.from_cplx(out)

    def z_collimator(self, n, outer, inner):
        out = np.zeros(n, self.cplx)
        self.lib.or_z_collimator(n, outer, inner, _ptr(out))
        return self.from_cplx(out)

    def z_parallel_plates(self, n, f0, f_max, g):
        out = np.zeros(n, self.cplx)
        terms = self.lib.or_z_parallel_plates(n, f0, f_max, g, _ptr(out))
        return self.from_cplx(out), int(terms)

    def airy(self, x):
        vals = [C.c_double() for _ in range(4)]
        ovf = self.lib.or_airy(float(x), *[C.byref(v) for v in vals])
        return tuple(v.value for v in vals), bool(ovf)

    # ---- driver --------------------------------------------------------
    def sim(self, cfg, init, Zwake=None, Zcsr=None):
        return OracleSim(self, cfg, init, Zwake, Zcsr)


class OracleSim:
    """The main loop of src/main.cpp:948-1053 on one mesh."""

    def __init__(self, orc, cfg, init, Zwake, Zcsr):
        self.o = orc
        c = orc.Cfg()
        for k, v in cfg.items():
            if k == "slip":
                arr = (orc.creal * 3)(*list(v) + [0.0] * (3 - len(v)))
                c.slip = arr
                c.nslip = len(v)
            else:
                setattr(c, k, v)
        self.cfg = c
        self.n = c.n
        init = orc.arr(init).reshape(self.n, self.n)
        zw = orc.to_cplx(Zwake) if Zwake is not None else None
        zc = orc.to_cplx(Zcsr) if Zcsr is not None else None
        self._keep = (zw, zc)
        self.h = orc.lib.or_sim_create(C.byref(c), _ptr(init), _ptr(zw),
                                       _ptr(zc))

    def run(self, nsteps, want_csr=False, want_profiles=False):
        csr = np.empty(nsteps, self.o.real) if want_csr else None
        prof = (np.empty((nsteps, self.n), self.o.real)
                if want_profiles else None)
        self.o.lib.or_sim_run(self.h, nsteps, _ptr(csr), _ptr(prof))
        return csr, prof

    def get(self):
        mesh = np.empty((self.n, self.n), self.o.real)
        proj = np.empty(self.n, self.o.real)
        wake = np.empty(self.n, self.o.real)
        self.o.lib.or_sim_get(self.h, _ptr(mesh), _ptr(proj), _ptr(wake))
        return mesh, proj, wake

    def set_mesh(self, mesh):
        mesh = self.o.arr(mesh)
        self.o.lib.or_sim_set_mesh(self.h, _ptr(mesh))

    def __del__(self):
        if getattr(self, "h", None):
            self.o.lib.or_sim_destroy(self.h)
            self.h = None
