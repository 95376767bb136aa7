"""ctypes binding of libreactiveTransport.so, the drop-in C ABI
(include/reactive_transport.h, include/reactive_gpu_hem.h).

The argument types mirror the reference's own ctypes bindings
(tools/validate_reactive_transport.py:130-146, tools/validate_impinging_n2o_recovery.py:65-71).
Loading fails loudly when the HIP library has not been built: there is no
host fallback for any entry point.
"""
import ctypes as C
import os

from . import LIBRARY_PATH
from .abi import (Config, DeviceProfile, HemProfile, MemoryV2, OptionsV2, OptionsV21, Profile, ProfileV21, Stats,
                  StepReport, Token)

P = C.c_void_p
D = C.POINTER(C.c_double)
_SIGNATURES = {
    "reactive_transport_create": (P, [C.c_int, C.POINTER(Config), P, P, P, P, P, P, C.c_char_p, C.c_size_t]),
    "reactive_transport_create_v2": (P, [C.c_int, C.POINTER(Config), C.POINTER(OptionsV2), P, P, P, P, P, P,
                                         C.c_char_p, C.c_size_t]),
    "reactive_transport_create_v21": (P, [C.c_int, C.POINTER(Config), C.POINTER(OptionsV21), P, P, P, P, P, P,
                                          C.c_char_p, C.c_size_t]),
    "reactive_transport_destroy": (None, [P]),
    "reactive_transport_error": (C.c_char_p, [P]),
    "reactive_transport_is_cuda": (C.c_int, [P]),
    "reactive_transport_stats": (C.c_int, [P, C.POINTER(Stats)]),
    "reactive_transport_stable_step": (C.c_int, [P, P, P, C.c_double, C.c_double, D]),
    "reactive_transport_stable_step_primitives": (C.c_int, [P, P, P, C.c_double, C.c_double, D]),
    "reactive_transport_rhs": (C.c_int, [P, P, P, P, P, P, P]),
    "reactive_transport_stage": (C.c_int, [P, P, P, P, P, C.c_double, C.c_int, P]),
    "reactive_transport_upload_conserved": (C.c_int, [P, P, C.c_uint64]),
    "reactive_transport_stage_resident": (C.c_int, [P, P, P, P, C.c_double, C.c_int, C.c_uint64, C.c_uint64, P, P]),
    "reactive_transport_begin_attempt": (C.c_int, [P, C.c_char_p, C.c_uint64]),
    "reactive_transport_end_attempt": (C.c_int, [P, C.c_uint64, C.c_int]),
    "reactive_transport_advance_resident_v2": (C.c_int, [P, Token, Token, P, P, P, P, C.c_double, P]),
    "reactive_transport_download_conserved": (C.c_int, [P, Token, P]),
    "reactive_transport_memory_v2": (C.c_int, [P, C.POINTER(MemoryV2)]),
    "reactive_transport_profile": (C.c_int, [P, C.POINTER(Profile)]),
    "reactive_transport_device_profile": (C.c_int, [P, C.POINTER(DeviceProfile)]),
    "reactive_transport_profile_v21": (C.c_int, [P, C.POINTER(ProfileV21)]),
    "reactive_gpu_hem_numerical_policy_v1": (C.c_char_p, []),
    "reactive_gpu_hem_create_v1": (P, [P, C.c_size_t, C.c_size_t, C.c_char_p, C.c_size_t]),
    "reactive_gpu_hem_create_resident_v1": (P, [P, C.c_size_t, C.c_size_t, C.c_char_p, C.c_size_t]),
    "reactive_gpu_hem_device_bytes_v1": (C.c_size_t, [P]),
    "reactive_gpu_hem_destroy_v1": (None, [P]),
    "reactive_gpu_hem_run_v1": (C.c_int, [P, P, P, C.c_size_t, P, P, C.POINTER(HemProfile), C.c_char_p, C.c_size_t]),
    "reactive_gpu_hem_run_v2": (C.c_int, [P, P, P, P, C.c_size_t, P, P, C.POINTER(HemProfile), C.c_char_p,
                                          C.c_size_t]),
    "reactive_gpu_hem_run_resident_v1": (C.c_int, [P, P, P, P, P, P, C.c_size_t, P, C.POINTER(HemProfile),
                                                   C.c_char_p, C.c_size_t]),
    "reactive_gpu_hem_run_resident_noalloc_v1": (C.c_int, [P, P, P, P, P, P, C.c_size_t, P, C.POINTER(HemProfile),
                                                           C.c_char_p, C.c_size_t]),
    "reactive_amd_transport_set_flux": (C.c_int, [P, C.c_int]),
    "reactive_amd_hem_set_arithmetic": (C.c_int, [P, C.c_int]),
    "reactive_amd_hem_mixed_report": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "reactive_amd_step_create": (P, [P, P, C.c_char_p, C.c_size_t]),
    "reactive_amd_step_load": (C.c_int, [P, P, P]),
    "reactive_amd_step_run": (C.c_int, [P, C.c_double, C.c_double, C.c_int, C.c_int, C.POINTER(StepReport)]),
    "reactive_amd_step_read": (C.c_int, [P, P, P]),
    "reactive_amd_step_create_batch": (P, [C.POINTER(C.c_void_p), C.c_int, P, C.c_char_p, C.c_size_t]),
    "reactive_amd_step_report_box": (C.c_int, [P, C.c_int, C.POINTER(StepReport)]),
    "reactive_amd_flash_adjoint": (C.c_int, [P, C.POINTER(C.c_void_p), C.c_double, P, C.c_int, P, P, C.c_size_t, P,
                                             P, P, P, C.POINTER(C.c_uint64), C.c_char_p, C.c_size_t]),
    "reactive_amd_transport_stage_adjoint": (C.c_int, [P, P, P, P, P, P, P, P]),
    "reactive_amd_step_error": (C.c_char_p, [P]),
    "reactive_amd_step_destroy": (None, [P]),
}

# exported stubs with scalar arguments
_SIGNATURES.update({
    "reactive_transport_stage_resident_gas": (C.c_int, [P, P, P, C.c_double, C.c_int, C.c_uint64, C.c_uint64, P, P]),
    "reactive_transport_replace_stage_state_v1": (C.c_int, [P, Token, C.c_uint64, P]),
    "reactive_transport_capillary_resident_begin_v1": (C.c_int, [P, C.c_size_t, P, P, P, P, P, C.c_double, C.c_int,
                                                                 P]),
    "reactive_transport_capillary_resident_select_v1": (C.c_int, [P, C.c_int, P, P]),
    "reactive_transport_capillary_resident_update_v1": (C.c_int, [P, C.c_size_t, P, P]),
    "reactive_transport_capillary_resident_commit_v1": (C.c_int, [P, P, P, P, P]),
    "reactive_transport_implicit_coefficients_v1": (C.c_int, [P, P, P, C.c_size_t, P]),
    "reactive_transport_geometric_diagnostic_v1": (C.c_int, [P] * 10),
    "reactive_transport_wale_scalar_fields_v1": (C.c_int, [P] * 5),
    "reactive_transport_wale_pr_properties_v1": (C.c_int, [P, P, P, C.c_size_t, P]),
    "reactive_transport_wale_pr_properties_v2": (C.c_int, [P, P, P, C.c_size_t, P, C.c_int]),
    "reactive_closure_scalar_create_v1": (P, [C.c_size_t, C.c_char_p, C.c_size_t]),
    "reactive_closure_scalar_destroy_v1": (None, [P]),
    "reactive_closure_scalar_run_v1": (C.c_int, [P, P, C.c_size_t, P, P, C.c_char_p, C.c_size_t]),
    "reactive_gpu_hem_initialize_tp_v1": (C.c_int, [P, P, P, P, C.c_size_t, P, P, C.POINTER(HemProfile), C.c_char_p,
                                                    C.c_size_t]),
})

# entry points (and exported stubs) whose arguments are all addresses; all take the handle first
for _name, _n in (("reactive_transport_set_gas_thermo", 7),
                  ("reactive_transport_rhs_gas", 6), ("reactive_transport_gas_properties_resident", 5),
                  ("reactive_transport_set_capillary_v1", 2), ("reactive_transport_capillary_geometry_v1", 6),
                  ("reactive_transport_capillary_profile_v1", 2),
                  ("reactive_transport_capillary_resident_relax_v1", 1),
                  ("reactive_transport_install_cartesian_v1", 2), ("reactive_transport_cartesian_info_v1", 2),
                  ("reactive_transport_set_implicit_v1", 2), ("reactive_transport_set_wale_v1", 2),
                  ("reactive_transport_wale_profile_v1", 2), ("reactive_transport_wale_primitives_v1", 4),
                  ("reactive_transport_set_wale_scalars_v1", 2), ("reactive_transport_wale_scalar_profile_v1", 2),
                  ("reactive_transport_set_wale_pr_model_v1", 2), ("reactive_transport_wale_pr_profile_v1", 2)):
    _SIGNATURES[_name] = (C.c_int, [P] * _n)

_LIB = None


def load(path=None):
    """Load the HIP library (once) and set every known signature."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    path = path or os.environ.get("REACTIVE_AMD_LIBRARY", LIBRARY_PATH)
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library {path} is not built: run spuma-multiphysics_amd/build.sh "
                           "(there is no host fallback)")
    lib = C.CDLL(path)
    for name, (restype, argtypes) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _LIB = lib
    return lib


def ptr(a):
    """numpy array -> void* (None for None)."""
    return None if a is None else a.ctypes.data_as(C.c_void_p)
