"""ctypes binding of libmsd.so (include/msd_abi.h).

Loading the library does not touch the GPU; the first object created calls
msd_init(0).  There is no CPU fallback: without libmsd.so every constructor
raises.
"""
import ctypes as C
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmsd.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "msd_abi.h")

MSD_OK = 0
MSD_EINVAL = -1
MSD_EDEVICE = -2
MSD_EUNSUPPORTED = -3
MSD_EMISMATCH = -4
MSD_ENOMEM = -5

_lib = None


class MagLuaError(RuntimeError):
    """An msd_* call failed; the message is msd_last_error() (MagLua raises luaL_error)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def declared_symbols(header=HEADER):
    """Names of the msd_* functions declared in include/msd_abi.h."""
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(msd_[a-z0-9_]+)\s*\(", txt)))


_vp = C.c_void_p
_pp = C.POINTER(C.c_void_p)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_lp = C.POINTER(C.c_long)

_SIGS = {
    "msd_last_error": (C.c_char_p, []),
    "msd_version": (C.c_int, []),
    "msd_init": (C.c_int, [C.c_int]),
    "msd_synchronize": (C.c_int, []),
    "msd_philox_device": (C.c_int, [_u32p, C.c_int, _u32p, _u32p]),
    "msd_system_create": (C.c_int, [C.c_int, C.c_int, C.c_int, _pp]),
    "msd_system_destroy": (C.c_int, [_vp]),
    "msd_system_dims": (C.c_int, [_vp, _ip, _ip, _ip]),
    "msd_system_copy": (C.c_int, [_vp, _pp]),
    "msd_system_copy_spins_to": (C.c_int, [_vp, _vp]),
    "msd_system_set_spins": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "msd_system_get_spins": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "msd_system_set_spin": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]),
    "msd_system_get_spin": (C.c_int, [_vp, C.c_int, _dp]),
    "msd_system_random_spins": (C.c_int, [_vp, C.c_uint64]),
    "msd_system_set_param": (C.c_int, [_vp, C.c_int, C.c_double]),
    "msd_system_get_param": (C.c_int, [_vp, C.c_int, _dp]),
    "msd_system_set_site_alpha": (C.c_int, [_vp, _dp]),
    "msd_hash32": (C.c_int, [C.c_char_p, C.POINTER(C.c_int32)]),
    "msd_system_encode": (C.c_int, [_vp, C.c_int, _vp, C.c_int64, C.POINTER(C.c_int64), _ip]),
    "msd_system_decode": (C.c_int, [_vp, C.c_int64, C.c_int, _pp, C.POINTER(C.c_int64), _ip]),
    "msd_exchange_encode": (C.c_int, [_vp, C.c_int, _vp, C.c_int64, C.POINTER(C.c_int64), _ip]),
    "msd_exchange_decode": (C.c_int, [_vp, C.c_int64, C.c_int, _pp, C.POINTER(C.c_int64), _ip]),
    "msd_anisotropy_encode": (C.c_int, [_vp, C.c_int, _vp, C.c_int64, C.POINTER(C.c_int64), _ip]),
    "msd_anisotropy_decode": (C.c_int, [_vp, C.c_int64, C.c_int, _pp, C.POINTER(C.c_int64), _ip]),
    "msd_dipole_encode": (C.c_int, [_vp, C.c_int, _vp, C.c_int64, C.POINTER(C.c_int64), _ip]),
    "msd_dipole_decode": (C.c_int, [_vp, C.c_int64, C.c_int, _pp, C.POINTER(C.c_int64), _ip]),
    "msd_system_set_site_gamma": (C.c_int, [_vp, _dp]),
    "msd_system_reset_fields": (C.c_int, [_vp]),
    "msd_system_sum_fields": (C.c_int, [_vp]),
    "msd_system_get_field": (C.c_int, [_vp, C.c_char_p, _dp, _dp, _dp]),
    "msd_system_slot_used": (C.c_int, [_vp, C.c_char_p, _ip]),
    "msd_system_net_moment": (C.c_int, [_vp, _dp]),
    "msd_net_moment_noalloc": (C.c_int, [_vp, _vp, C.c_int, _dp]),
    "msd_system_field_dot": (C.c_int, [_vp, C.c_char_p, _dp]),
    "msd_system_diff": (C.c_int, [_vp, _vp, _dp]),
    "msd_exchange_create": (C.c_int, [C.c_int, C.c_int, C.c_int, _pp]),
    "msd_exchange_destroy": (C.c_int, [_vp]),
    "msd_exchange_set_pbc": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "msd_exchange_set_normalize": (C.c_int, [_vp, C.c_int]),
    "msd_exchange_set_scale": (C.c_int, [_vp, C.c_double]),
    "msd_exchange_add": (C.c_int, [_vp, _ip, _ip, C.c_double, _ip]),
    "msd_exchange_add_path": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double]),
    "msd_exchange_add_paths": (C.c_int, [_vp, _ip, _ip, _dp, C.c_int]),
    "msd_exchange_num_paths": (C.c_int, [_vp, _ip]),
    "msd_exchange_apply": (C.c_int, [_vp, _vp]),
    "msd_anisotropy_create": (C.c_int, [C.c_int, C.c_int, C.c_int, _pp]),
    "msd_anisotropy_destroy": (C.c_int, [_vp]),
    "msd_anisotropy_set_scale": (C.c_int, [_vp, C.c_double]),
    "msd_anisotropy_add": (C.c_int, [_vp, C.c_int, _dp, C.c_double, _ip]),
    "msd_anisotropy_add_many": (C.c_int, [_vp, _ip, _dp, _dp, C.c_int]),
    "msd_anisotropy_apply": (C.c_int, [_vp, _vp]),
    "msd_uniaxial_create": (C.c_int, [C.c_int, C.c_int, C.c_int, _pp]),
    "msd_uniaxial_destroy": (C.c_int, [_vp]),
    "msd_uniaxial_set_scale": (C.c_int, [_vp, C.c_double]),
    "msd_uniaxial_add": (C.c_int, [_vp, C.c_int, _dp, C.c_double, C.c_double, _ip]),
    "msd_uniaxial_num_axes": (C.c_int, [_vp, _ip]),
    "msd_uniaxial_axis_at": (C.c_int, [_vp, C.c_int, _ip, _dp, _dp, _dp]),
    "msd_uniaxial_apply": (C.c_int, [_vp, _vp]),
    "msd_appliedfield_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _pp]),
    "msd_appliedfield_destroy": (C.c_int, [_vp]),
    "msd_appliedfield_set": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double]),
    "msd_appliedfield_get": (C.c_int, [_vp, _dp]),
    "msd_appliedfield_set_scale": (C.c_int, [_vp, C.c_double]),
    "msd_appliedfield_set_site": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_double]),
    "msd_appliedfield_get_site": (C.c_int, [_vp, C.c_int, _dp]),
    "msd_appliedfield_apply": (C.c_int, [_vp, _vp]),
    "msd_stepper_add_applied_field": (C.c_int, [_vp, _vp]),
    "msd_stepper_set_uniaxial": (C.c_int, [_vp, _vp]),
    "msd_thermal_create": (C.c_int, [C.c_int, C.c_int, C.c_int, _pp]),
    "msd_thermal_destroy": (C.c_int, [_vp]),
    "msd_thermal_set_temperature": (C.c_int, [_vp, C.c_double]),
    "msd_thermal_set_scale": (C.c_int, [_vp, C.c_double]),
    "msd_thermal_scale_site": (C.c_int, [_vp, C.c_int, C.c_double]),
    "msd_thermal_set_seed": (C.c_int, [_vp, C.c_uint64, C.c_uint64]),
    "msd_thermal_encode": (C.c_int, [_vp, C.c_int, _vp, C.c_int64, C.POINTER(C.c_int64), _ip]),
    "msd_uniaxial_encode": (C.c_int, [_vp, C.c_int, _vp, C.c_int64, C.POINTER(C.c_int64), _ip]),
    "msd_uniaxial_decode": (C.c_int, [_vp, C.c_int64, C.c_int, _pp, C.POINTER(C.c_int64), _ip]),
    "msd_appliedfield_encode": (C.c_int, [_vp, C.c_int, _vp, C.c_int64, C.POINTER(C.c_int64), _ip]),
    "msd_appliedfield_decode": (C.c_int, [_vp, C.c_int64, C.c_int, C.c_int, _pp, C.POINTER(C.c_int64), _ip]),
    "msd_thermal_decode": (C.c_int, [_vp, C.c_int64, C.c_int, _pp, C.POINTER(C.c_int64), _ip]),
    "msd_thermal_get_draw": (C.c_int, [_vp, _u64p]),
    "msd_thermal_get_seed": (C.c_int, [_vp, _u64p]),
    "msd_thermal_apply": (C.c_int, [_vp, _vp]),
    "msd_dipole_create": (C.c_int, [C.c_int, C.c_int, C.c_int, _pp]),
    "msd_dipole_destroy": (C.c_int, [_vp]),
    "msd_dipole_set_strength": (C.c_int, [_vp, C.c_double]),
    "msd_dipole_set_scale": (C.c_int, [_vp, C.c_double]),
    "msd_dipole_set_truncation": (C.c_int, [_vp, _ip]),
    "msd_dipole_schedule": (C.c_int, [_vp, _ip, _ip, _ip]),
    "msd_dipole_set_unit_cell": (C.c_int, [_vp, _dp]),
    "msd_dipole_get_matrix": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
    "msd_dipole_set_matrix": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]),
    "msd_dipole_apply": (C.c_int, [_vp, _vp]),
    "msd_lr2d_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_char_p, _pp]),
    "msd_lr2d_set_tensor": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp]),
    "msd_lr2d_get_tensor": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp]),
    "msd_lr2d_set_layer_scales": (C.c_int, [_vp, _dp, _dp]),
    "msd_lr2d_get_layer_scales": (C.c_int, [_vp, _dp, _dp]),
    "msd_lr2d_build_dipole2d": (C.c_int, [_vp, _dp, _dp, _ip, C.c_int]),
    "msd_lr2d_build_magnetostatics2d": (C.c_int, [_vp, _dp, _dp, _ip]),
    "msd_longrange3d_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _pp]),
    "msd_longrange3d_destroy": (C.c_int, [_vp]),
    "msd_longrange3d_set_strength": (C.c_int, [_vp, C.c_double]),
    "msd_longrange3d_set_scale": (C.c_int, [_vp, C.c_double]),
    "msd_longrange3d_set_radius": (C.c_int, [_vp, C.c_int]),
    "msd_longrange3d_get_matrix": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
    "msd_longrange3d_set_matrix": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]),
    "msd_longrange3d_apply": (C.c_int, [_vp, _vp]),
    "msd_longrange3d_set_pass_widths": (C.c_int, [_vp, C.c_int, C.c_int]),
    "msd_longrange3d_pass_widths": (C.c_int, [_vp, _ip, _ip, _ip, _ip, _ip, _ip]),
    "msd_stepper_set_longrange3d": (C.c_int, [_vp, _vp]),
    "msd_ensemble_create": (C.c_int, [_pp, C.c_int, _pp]),
    "msd_ensemble_destroy": (C.c_int, [_vp]),
    "msd_ensemble_set_graph": (C.c_int, [_vp, C.c_int]),
    "msd_ensemble_set_compensated": (C.c_int, [_vp, C.c_int]),
    "msd_ensemble_heun": (C.c_int, [_vp, C.c_int]),
    "msd_ensemble_profile": (C.c_int, [_vp, C.c_int, C.c_int, C.c_void_p, _dp, _ip]),
    "msd_llg_create": (C.c_int, [_pp]),
    "msd_llg_create_quaternion": (C.c_int, [_pp]),
    "msd_llg_destroy": (C.c_int, [_vp]),
    "msd_llg_set_thermal_only_first_term": (C.c_int, [_vp, C.c_int]),
    "msd_llg_set_disable_renormalization": (C.c_int, [_vp, C.c_int]),
    "msd_llg_set_scales": (C.c_int, [_vp, C.c_double, C.c_double]),
    "msd_llg_encode": (C.c_int, [_vp, C.c_int, _vp, C.c_int64, C.POINTER(C.c_int64), _ip]),
    "msd_llg_decode": (C.c_int, [_vp, C.c_int64, C.c_int, C.c_int, _pp, C.POINTER(C.c_int64), _ip]),
    "msd_llg_apply": (C.c_int, [_vp, _vp, C.c_double, _vp, _vp, C.c_int]),
    "msd_stepper_create": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _pp]),
    "msd_stepper_destroy": (C.c_int, [_vp]),
    "msd_stepper_set_graph": (C.c_int, [_vp, C.c_int]),
    "msd_stepper_set_compensated": (C.c_int, [_vp, C.c_int]),
    "msd_stepper_heun": (C.c_int, [_vp, C.c_int]),
    "msd_stepper_profile": (C.c_int, [_vp, C.c_int, C.c_int, C.c_void_p, _dp, _ip]),
    "msd_stepper_set_surrogate": (C.c_int, [_vp, _vp, C.c_int, C.c_double]),
    "msd_stepper_autotune": (C.c_int, [_vp, C.c_int, _ip, _ip, _ip, _dp]),
    "msd_stepper_autotune_passes": (C.c_int, [_vp, C.c_int, _ip, _ip, _dp]),
    "msd_ensemble_autotune_passes": (C.c_int, [_vp, C.c_int, _dp]),
    "msd_stepper_rk": (C.c_int, [_vp, C.c_char_p, C.c_int]),
    "msd_stepper_graph_builds": (C.c_int, [_vp, _lp]),
    "msd_stepper_surrogate_stats": (C.c_int, [_vp, _lp, _lp, _lp, _dp, _dp]),
    "msd_surrogate_create": (C.c_int, [C.c_int, C.c_int, C.c_int, _pp]),
    "msd_surrogate_destroy": (C.c_int, [_vp]),
    "msd_surrogate_set_open": (C.c_int, [_vp, C.c_int]),
    "msd_surrogate_set_weights": (C.c_int, [_vp, C.c_int, _dp]),
    "msd_surrogate_get_weights": (C.c_int, [_vp, C.c_int, _dp]),
    "msd_surrogate_set_scales": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double]),
    "msd_surrogate_get_scales": (C.c_int, [_vp, _dp, _dp, _dp]),
    "msd_surrogate_set_input": (C.c_int, [_vp, _dp]),
    "msd_surrogate_set_input_systems": (C.c_int, [_vp, _vp, _vp]),
    "msd_surrogate_set_target": (C.c_int, [_vp, _dp]),
    "msd_surrogate_set_target_dipole": (C.c_int, [_vp, _vp, _vp, _vp]),
    "msd_surrogate_set_target_longrange3d": (C.c_int, [_vp, _vp, _vp, _vp]),
    "msd_surrogate_forward": (C.c_int, [_vp, _dp]),
    "msd_surrogate_loss_grads": (C.c_int, [_vp, _dp]),
    "msd_surrogate_get_grads": (C.c_int, [_vp, C.c_int, _dp]),
    "msd_surrogate_adam": (C.c_int, [_vp, C.c_double]),
    "msd_surrogate_reset_optimizer": (C.c_int, [_vp]),
    "msd_surrogate_train_step": (C.c_int, [_vp, C.c_double, _dp]),
    "msd_surrogate_field_error": (C.c_int, [_vp, _vp, _vp, _vp, _dp, _dp]),
    "msd_surrogate_field_error_longrange3d": (C.c_int, [_vp, _vp, _vp, _vp, _dp, _dp]),
}


def lib():
    """The loaded libmsd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built (run __graft_entry__.build() or make -C maglua_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != MSD_OK:
        raise MagLuaError(rc, lib().msd_last_error().decode(errors="replace"))
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args))
