"""ctypes binding of libgpukern.so (include/gpukern.h + include/mumax_engine.h).

The library is the product; this module only loads it and declares argument
types.  It raises immediately if the library has not been built: there is no
fallback implementation of any entry point.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpukern.so")
# launch parameters chosen by the autotuner on MI355X (1_amd/tune.py), read at load
TUNE_TABLE = os.path.join(HERE, "tune_gfx950.txt")

P = C.c_void_p
F = C.c_float
I = C.c_int
FP = C.POINTER(C.c_float)
IP = C.POINTER(C.c_int)

# name: (restype, argtypes) for every entry point of include/gpukern.h
GPUKERN = {
    "gpu_init": (None, [I, I]),
    "gpu_set_device": (None, [I]),
    "gpu_maxthreads": (I, []),
    "gpu_print_properties_stdout": (None, []),
    "new_gpu_array": (P, [I]),
    "free_gpu_array": (None, [P]),
    "gpu_usedmem": (C.c_ulonglong, []),
    "gpu_stride_float": (I, []),
    "gpu_override_stride": (None, [I]),
    "memcpy_gpu_dir": (None, [P, P, I, I]),
    "memcpy_to_gpu": (None, [P, P, I]),
    "memcpy_from_gpu": (None, [P, P, I]),
    "memcpy_on_gpu": (None, [P, P, I]),
    "gpu_array_get": (F, [P, I]),
    "gpu_array_set": (None, [P, I, F]),
    "gpu_zero": (None, [P, I]),
    "gpu_add": (None, [P, P, I]),
    "gpu_madd": (None, [P, F, P, I]),
    "gpu_madd2": (None, [P, P, P, I]),
    "gpu_add_constant": (None, [P, F, I]),
    "gpu_linear_combination": (None, [P, P, F, F, I]),
    "gpu_linear_combination_many": (None, [P, P, P, I, I]),
    "gpu_scale_dot_product": (None, [P, P, P, F, I]),
    "gpu_reduce": (F, [I, P, P, P, I, I, I]),
    "gpu_normalize_uniform": (None, [P, I]),
    "gpu_normalize_map": (None, [P, P, I]),
    "gpu_deltaM": (None, [P, P, F, P, F, I]),
    "gpu_spintorque_deltaM": (None, [P, P, F, F, F, P, P, F, I, I, I]),
    "gpu_add_local_fields": (None, [P, P, I, P, P, I, P, P]),
    "gpu_add_local_fields_H_and_phi": (None, [P, P, P, I, P, P, I, P, P]),
    "gpu_add_lin_anisotropy": (None, [P] * 12 + [I]),
    "gpu_add_exch": (None, [P, P, P, P, P, P, I]),
    "gpu_anal_fw_step": (None, [F, F, I, P, P, P]),
    "gpu_kernelmul6": (None, [P] * 9 + [I]),
    "gpu_kernelmul4": (None, [P] * 7 + [I]),
    "gpu_copy_pad": (None, [P, P, I, I, I, I, I, I]),
    "gpu_copy_unpad": (None, [P, P, I, I, I, I, I, I]),
    "new_gpuFFT3dPlan_padded": (P, [P, P]),
    "delete_gpuFFT3dPlan": (None, [P]),
    "gpuFFT3dPlan_forward": (None, [P, P, P]),
    "gpuFFT3dPlan_forward_noalloc": (None, [P, P, P]),
    "gpuFFT3dPlan_inverse": (None, [P, P, P]),
    "gpuFFT3dPlan_normalization": (I, [P]),
    "gpu_gaussian_noise": (None, [P, F, F, I]),
}

CP = C.c_char_p

ENGINE = {
    "mm_sim_new": (P, [P, P, I]),
    "mm_sim_free": (None, [P]),
    "mm_sim_run": (C.c_long, [P, C.c_double]),
    "mm_sim_steps": (None, [P, I]),
    "mm_sim_steps_async": (None, [P, I]),
    "mm_sim_sync": (None, [P]),
    "mm_sim_get_m": (None, [P, P]),
    "mm_sim_set_m": (None, [P, P]),
    "mm_sim_device_m": (P, [P]),
    "mm_sim_state": (None, [P, P]),
    "mm_sim_trial_log": (I, [P, P, P, I]),
    "mm_sim_trial_errors": (I, [P, P, I]),
    "mm_sim_heff": (None, [P, P, P]),
    "mm_sim_torque": (None, [P, P, P]),
    "mm_sim_energy": (C.c_double, [P]),
    "mm_sim_average_m": (None, [P, P]),
    "mm_sim_kernel_spectrum": (C.c_long, [P, P]),
    "mm_sim_device_bytes": (C.c_ulonglong, [P]),
    "mm_sim_time_steps": (C.c_double, [P, I]),
    "mm_sim_profile_steps": (C.c_double, [P, I, P, P]),
    "mm_sim_time_pass": (C.c_double, [P, I, I]),
    "mm_sim_tune": (I, [P, I, P, P, P]),
    "mm_sim_get_tune": (None, [P, P]),
    "mm_tune_save": (I, [CP]),
    "mm_tune_load": (I, [CP]),
    "mm_tune_clear": (None, []),
    "mm_sim_footprint": (C.c_ulonglong, [P]),
    "mm_sim_begin_run": (None, [P, C.c_double]),
    "mm_sim_relax": (C.c_long, [P, C.c_long]),
    "mm_sim_run_until": (C.c_long, [P, C.c_double]),
    "mm_demag_kernel": (None, [P, P, I, I, P]),
    "mm_kernel_cache_size": (I, []),
    "mm_memory_cached": (C.c_ulonglong, []),
    "mm_memory_trim": (C.c_ulonglong, []),
    "mm_sim_set_field": (None, [P, I, P, P]),
    "mm_sim_set_field_function": (None, [P, I, P, P]),
    "mm_sim_set_temperature": (None, [P, C.c_double]),
    "mm_sim_set_spintorque": (None, [P, P, C.c_double, C.c_double]),
    "mm_sim_set_current": (None, [P, I, P, P, I, P]),
    "mm_batch_new": (P, [P, I]),
    "mm_batch_free": (None, [P]),
    "mm_batch_set_graph": (None, [P, I]),
    "mm_batch_set_merge_blocks": (I, [I]),
    "mm_batch_size": (I, [P]),
    "mm_batch_start": (None, [P, P, P]),
    "mm_batch_trials": (I, [P, I]),
    "mm_batch_enqueue": (None, [P, I]),
    "mm_batch_ready": (I, [P]),
    "mm_batch_flags": (I, [P, P]),
    "mm_batch_remaining": (I, [P, P]),
    "mm_batch_done": (None, [P, P]),
    "mm_batch_run": (C.c_longlong, [P, P, I]),
}


class Params(C.Structure):
    """mmParams (include/mumax_engine.h)."""

    _fields_ = [
        ("size", C.c_int * 3),
        ("periodic", C.c_int * 3),
        ("cellsize", C.c_double * 3),
        ("msat", C.c_double),
        ("aexch", C.c_double),
        ("alpha", C.c_double),
        ("anisType", C.c_int),
        ("k1", C.c_double),
        ("anisAxis", C.c_double * 3),
        ("exchType", C.c_int),
        ("exchInConv", C.c_int),
        ("wantDemag", C.c_int),
        ("demagAccuracy", C.c_int),
        ("solver", C.c_int),
        ("maxError", C.c_double),
        ("dt", C.c_double),
        ("minDm", C.c_double),
        ("maxDm", C.c_double),
        ("minDt", C.c_double),
        ("maxDt", C.c_double),
        ("nField", C.c_int),
        ("fieldT", C.POINTER(C.c_double)),
        ("fieldB", C.POINTER(C.c_double)),
        ("wantEnergy", C.c_int),
        ("demagKernel", C.c_int),
        ("alphaMap", C.POINTER(C.c_float)),
        ("temperature", C.c_double),
        ("seed", C.c_ulonglong),
        ("normMap", C.POINTER(C.c_float)),
        ("current", C.c_double * 3),
        ("spinPol", C.c_double),
        ("xi", C.c_double),
        ("currentMask", C.POINTER(C.c_float)),
    ]


_lib = None


PP = C.POINTER(C.c_char_p)

IO = {
    "mm_io_error": (CP, []),
    "mm_ovf_info": (I, [CP, P]),
    "mm_ovf_desc": (I, [CP, I, P, P, I]),
    "mm_ovf_read": (I, [CP, P, P]),
    "mm_ovf_write": (I, [CP, P, P, PP, PP, I]),
    "mm_tensor_read": (I, [CP, P, P, P]),
    "mm_tensor_write": (I, [CP, I, P, P]),
    "mm_odt_open": (P, [CP, CP, I, PP, PP]),
    "mm_odt_open_desc": (P, [CP, CP, I, PP, PP, I, PP, PP]),
    "mm_odt_row": (I, [P, P]),
    "mm_odt_close": (I, [P]),
}


def load(path=LIB_PATH):
    """Load libgpukern.so (once) and declare every entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            "libgpukern.so is not built (%s); run `python 1_amd/build.py` or __graft_entry__.build()" % path)
    lib = C.CDLL(path)
    for table in (GPUKERN, ENGINE, IO):
        for name, (res, args) in table.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
    _lib = lib
    if os.path.exists(TUNE_TABLE):
        lib.mm_tune_load(TUNE_TABLE.encode())
    return lib


def exported_symbols(path=LIB_PATH):
    """Names from both tables that the shared object exports (no GPU needed)."""
    lib = C.CDLL(path)
    return {n for n in list(GPUKERN) + list(ENGINE) + list(IO) if hasattr(lib, n)}
