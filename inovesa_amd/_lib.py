"""ctypes binding of libvfps.so (the C ABI in include/vfps_gpu.h).

The library is built in-tree by inovesa_amd/build.py (hipcc, gfx950).  There is
no CPU fallback: if the shared object is missing, importing this module
raises, and every call that fails on the device raises VfpsError with the
library's message.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libvfps.so")


class VfpsError(RuntimeError):
    """A non-zero status from the C ABI."""

    def __init__(self, code, msg):
        super().__init__(f"vfps error {code}: {msg}")
        self.code = code


class VfpsInvalid(VfpsError, ValueError):
    """VFPS_ERR_INVALID: the reference throws std::invalid_argument here."""


if not os.path.exists(LIBPATH):
    raise ImportError(
        f"{LIBPATH} is missing: build it with `python inovesa_amd/build.py` "
        "(hipcc, gfx950).  There is no CPU fallback.")

lib = C.CDLL(LIBPATH)

_u32 = C.c_uint32
_i32 = C.c_int32
_f32 = C.c_float
_f64 = C.c_double
_vp = C.c_void_p
_int = C.c_int
_sz = C.c_size_t


class RingParams(C.Structure):
    _fields_ = [("grid_size", _u32), ("pq_size", _f64), ("shift_x", _f64),
                ("shift_y", _f64), ("energy_spread", _f64),
                ("beam_energy", _f64), ("bending_radius", _f64),
                ("f_rev", _f64), ("harmonic_number", _f64),
                ("rf_voltage", _f64), ("f_s", _f64), ("alpha0", _f64),
                ("alpha1", _f64), ("alpha2", _f64), ("vacuum_gap", _f64),
                ("damping_time", _f64), ("padding", _f64),
                ("round_padding", _i32), ("steps_per_ts", _u32),
                ("rotations", _f64), ("bunch_current", _f64),
                ("cutoff_frequency", _f64), ("steps_per_trev", _f64)]


class Derived(C.Structure):
    _fields_ = [("qmin", _f32), ("qmax", _f32), ("pmin", _f32),
                ("pmax", _f32), ("qscale", _f64), ("pscale", _f64),
                ("R_bend", _f64), ("f0", _f64), ("V0", _f64), ("V_eff", _f64),
                ("f_s", _f64), ("alpha0", _f64), ("f_RF", _f64),
                ("steps", _f64), ("dt", _f64), ("revolutionpart", _f64),
                ("angle", _f32), ("slip", _f32 * 3), ("e1", _f32),
                ("t_damp", _f64), ("padded_bins", _u32),
                ("spacing_bins", _u32), ("fmax", _f32), ("Ib", _f64),
                ("wakescaling", _f32), ("shield", _f64), ("Inorm", _f64),
                ("Ith", _f64), ("laststep", _u32)]


class StepDesc(C.Structure):
    _fields_ = [("n", _u32), ("qmin", _f32), ("qmax", _f32), ("pmin", _f32),
                ("pmax", _f32), ("qscale", _f64), ("pscale", _f64),
                ("scheme", _u32), ("interp", _u32), ("rot_clamp", _i32),
                ("angle", _f32), ("rf_linear", _i32), ("rf_angle", _f32),
                ("revolutionpart", _f32), ("V_RF", _f32), ("f_RF", _f32),
                ("V0", _f32), ("slip", _f32 * 3), ("nslip", _u32),
                ("E0", _f32), ("use_fp", _i32), ("fptype", _u32),
                ("e1", _f32), ("deriv", _u32), ("use_wake", _i32),
                ("spacing_bins", _u32), ("wakescaling", _f32),
                ("use_csr", _i32), ("fc", _f32), ("fused", _i32),
                ("tile_rows", _u32), ("fast_projection", _i32),
                ("wake_method", _u32), ("threads", _u32), ("no_graph", _i32),
                ("renormalize", _u32)]


_SIGS = {
    "vfps_last_error": (C.c_char_p, []),
    "vfps_version": (C.c_char_p, []),
    "vfps_dev_open": (_int, [_int, _vp]),
    "vfps_dev_close": (_int, [_vp]),
    "vfps_sync": (_int, [_vp]),
    "vfps_mesh_create": (_int, [_vp, _u32, _u32, _f32, _f32, _f64, _f32, _f32,
                                _f64, _f64, _f64, _vp, _f64, _vp, _vp]),
    "vfps_mesh_clone": (_int, [_vp, _vp]),
    "vfps_mesh_destroy": (_int, [_vp]),
    "vfps_mesh_upload": (_int, [_vp, _vp]),
    "vfps_mesh_download": (_int, [_vp, _vp]),
    "vfps_mesh_copy": (_int, [_vp, _vp]),
    "vfps_mesh_device_ptr": (_int, [_vp, _vp]),
    "vfps_mesh_project_x": (_int, [_vp]),
    "vfps_mesh_project_y": (_int, [_vp]),
    "vfps_mesh_get_projection": (_int, [_vp, _int, _vp]),
    "vfps_mesh_set_projection": (_int, [_vp, _int, _vp]),
    "vfps_mesh_integrate": (_int, [_vp, _vp, _vp]),
    "vfps_mesh_normalize": (_int, [_vp]),
    "vfps_mesh_variance": (_int, [_vp, _int, _vp, _vp, _vp]),
    "vfps_mesh_info": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "vfps_mesh_axis": (_int, [_vp, _int, _vp]),
    "vfps_mesh_simpson_weights": (_int, [_vp, _vp]),
    "vfps_map_create_kick": (_int, [_vp, _vp, _int, _int, _int, _vp]),
    "vfps_map_set_offset": (_int, [_vp, _vp, _sz]),
    "vfps_map_get_offset": (_int, [_vp, _vp, _sz]),
    "vfps_map_create_drift": (_int, [_vp, _vp, _vp, _u32, _f32, _int, _int,
                                     _vp]),
    "vfps_map_create_rfkick_linear": (_int, [_vp, _vp, _f32, _f32, _int, _int,
                                             _vp]),
    "vfps_map_create_rfkick": (_int, [_vp, _vp, _f32, _f32, _f32, _f32, _int,
                                      _int, _vp]),
    "vfps_map_create_dynrfkick": (_int, [_vp, _vp, C.c_double, C.c_float,
                                         C.c_float, C.c_float, C.c_float,
                                         C.c_float, C.c_float, C.c_double,
                                         _u32, _u32, _int, _int, _vp]),
    "vfps_map_create_dynrfkick_linear": (_int, [_vp, _vp, C.c_float,
                                                C.c_double, C.c_float,
                                                C.c_float, C.c_float,
                                                C.c_float, C.c_double, _u32,
                                                _u32, _int, _int, _vp]),
    "vfps_map_get_modulation": (_int, [_vp, _u32, _u32, _vp, _vp]),
    "vfps_dynrf_modulation": (_int, [_u32, _u32, C.c_float, C.c_double,
                                     C.c_float, C.c_float, C.c_float,
                                     C.c_double, _vp]),
    "vfps_map_create_rotation": (_int, [_vp, _vp, _f32, _int, _int, _u32,
                                        _vp]),
    "vfps_map_create_fokkerplanck": (_int, [_vp, _vp, _int, _int, _f32, _int,
                                            _vp]),
    "vfps_map_create_identity": (_int, [_vp, _vp, _vp]),
    "vfps_map_create_wakepotential": (_int, [_vp, _vp, _vp, _int, _int, _vp]),
    "vfps_map_update": (_int, [_vp]),
    "vfps_map_apply": (_int, [_vp]),
    "vfps_map_destroy": (_int, [_vp]),
    "vfps_impedance_create": (_int, [_vp, _u32, _f32, _vp]),
    "vfps_impedance_create_table": (_int, [_vp, _vp, _u32, _f32, _vp]),
    "vfps_impedance_destroy": (_int, [_vp]),
    "vfps_impedance_add_freespace_csr": (_int, [_vp, _f32, _f32]),
    "vfps_impedance_add_parallel_plates": (_int, [_vp, _f32, _f32, _f64]),
    "vfps_impedance_add_resistive_wall": (_int, [_vp, _f32, _f32, _f64, _f64,
                                                 _f64, _f64]),
    "vfps_impedance_add_const": (_int, [_vp, _f32, _f32]),
    "vfps_impedance_add_collimator": (_int, [_vp, _f64, _f64]),
    "vfps_impedance_add": (_int, [_vp, _vp]),
    "vfps_impedance_download": (_int, [_vp, _vp]),
    "vfps_impedance_nfreqs": (_int, [_vp, _vp]),
    "vfps_airy": (_int, [_vp, _vp, _u32, _vp, _vp]),
    "vfps_wake_potential_noalloc": (_int, [_vp, _vp, _u32, _u32, _vp, _u32,
                                           _vp, _u32, C.c_float, _vp]),
    "vfps_field_get_padded_wake": (_int, [_vp, _vp]),
    "vfps_field_create": (_int, [_vp, _vp, _vp, _u32, _u32, _f64, _f32, _f32,
                                 _vp]),
    "vfps_field_wakescaling": (_f32, [_f64, _f64, _f64, _f32, _f64, _f64]),
    "vfps_field_destroy": (_int, [_vp]),
    "vfps_field_pad": (_int, [_vp]),
    "vfps_field_get_padded": (_int, [_vp, _vp]),
    "vfps_field_wake_potential": (_int, [_vp, _vp]),
    "vfps_field_get_wakescaling": (_int, [_vp, _vp]),
    "vfps_field_update_csr": (_int, [_vp, _f32, _vp, _vp]),
    "vfps_derive": (_int, [C.POINTER(RingParams), C.POINTER(Derived)]),
    "vfps_step_create": (_int, [_vp, C.POINTER(StepDesc), _vp, _vp, _vp,
                                _vp]),
    "vfps_step_destroy": (_int, [_vp]),
    "vfps_step_run": (_int, [_vp, _u32]),
    "vfps_step_prepare": (_int, [_vp, _u32]),
    "vfps_step_record": (_int, [_vp, _u32, _int]),
    "vfps_step_get_history": (_int, [_vp, _vp, _vp, _vp]),
    "vfps_step_get_mesh": (_int, [_vp, _vp, _vp, _vp]),
    "vfps_step_set_mesh": (_int, [_vp, _vp]),
    "vfps_step_moments": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "vfps_step_set_particles": (_int, [_vp, _vp, _u32, _int, C.c_uint64]),
    "vfps_step_get_particles": (_int, [_vp, _vp, _u32]),
    "vfps_map_apply_to": (_int, [_vp, _vp, _u32, C.c_uint64]),
    "vfps_step_timing": (_int, [_vp, _u32, _vp, _vp, _vp]),
    "vfps_step_phases": (_int, [_vp, _vp, _u32]),
    "vfps_step_config": (_int, [_vp, _vp, _u32]),
    "vfps_batch_create": (_int, [_vp, _u32, C.POINTER(StepDesc), _vp, _vp,
                                 _vp, _vp]),
    "vfps_batch_destroy": (_int, [_vp]),
    "vfps_batch_run": (_int, [_vp, _u32]),
    "vfps_batch_prepare": (_int, [_vp, _u32]),
    "vfps_batch_record": (_int, [_vp, _u32, _int]),
    "vfps_batch_get_history": (_int, [_vp, _u32, _vp, _vp, _vp]),
    "vfps_batch_get_mesh": (_int, [_vp, _u32, _vp, _vp, _vp]),
    "vfps_batch_info": (_int, [_vp, _vp, _vp, _vp]),
    "vfps_batch_get_projections": (_int, [_vp, _vp]),
    "vfps_batch_get_csr_histories": (_int, [_vp, _u32, _u32, _vp]),
    "vfps_batch_config": (_int, [_vp, _u32, _vp, _u32]),
    "vfps_batch_retile": (_int, [_vp, _u32, _u32, _u32, _i32]),
    "vfps_batch_launches": (_int, [_vp, _vp, _u32, _vp]),
    "vfps_batch_retile_size": (_int, [_vp, _u32, _u32, _u32, _i32]),
    "vfps_batch_retile_strip": (_int, [_vp, _u32, _u32, _u32, _u32]),
    "vfps_batch_retile_size_strip": (_int, [_vp, _u32, _u32, _u32, _u32]),
    "vfps_batch_set_merge": (_int, [_vp, _i32]),
    "vfps_batch_timing": (_int, [_vp, _u32, _vp]),
    "vfps_batch_autotune": (_int, [_vp, _u32, _vp, _u32, _vp]),
    "vfps_mlp_create": (_int, [_vp, _vp, _u32, _u32, C.c_uint64, _vp]),
    "vfps_mlp_destroy": (_int, [_vp]),
    "vfps_mlp_nparams": (_int, [_vp, _vp]),
    "vfps_mlp_set_params": (_int, [_vp, _vp, _sz]),
    "vfps_mlp_get_params": (_int, [_vp, _vp, _vp, _sz]),
    "vfps_mlp_forward": (_int, [_vp, _vp, _u32, _vp]),
    "vfps_mlp_loss_grad": (_int, [_vp, _vp, _vp, _vp, _u32, _f32, _vp]),
    "vfps_mlp_adam_step": (_int, [_vp, _f32, _f32, _f32, _f32]),
    "vfps_mlp_train": (_int, [_vp, _vp, _vp, _vp, _u32, _u32, _u32, _f32,
                              _f32, C.c_uint64, _vp]),
    "vfps_gemm_bf16": (_int, [_vp, _u32, _u32, _u32, _vp, _sz, C.c_int64,
                              C.c_int64, _vp, _sz, C.c_int64, C.c_int64, _vp]),
    "vfps_round_bf16": (_int, [_vp, _vp, _sz]),
    "vfps_toeplitz_kernels": (_int, [_u32, _u32, _vp, _vp, _vp, _f32, _vp,
                                     _vp]),
}

for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


def exported_symbols():
    """Names of the C ABI entry points declared in include/vfps_gpu.h."""
    return sorted(_SIGS)


def check(rc):
    """Raise on a non-zero status from the C ABI."""
    if rc != 0:
        msg = lib.vfps_last_error().decode(errors="replace")
        if rc == -1:
            raise VfpsInvalid(rc, msg)
        raise VfpsError(rc, msg)
    return rc


def call(name, *args):
    return check(getattr(lib, name)(*args))
