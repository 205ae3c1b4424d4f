"""The C-ABI library loads and exports every function include/*.h declares.

CPU only: nothing here launches a kernel.
"""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "1_amd", "libgpukern.so")

DECL = re.compile(r"^\s*(?:const\s+)?(?:unsigned\s+long\s+long\s+int|unsigned\s+long\s+long|long\s+long|[A-Za-z_][\w]*)\s*\*?\s*"
                  r"\**\s*([A-Za-z_]\w*)\s*\(", re.M)


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in DECL.finditer(text):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof", "defined"):
                names.add(name)
    return names


def test_headers_declare_reference_surface():
    names = declared_functions()
    # the 37 functions gpu.go binds (gpu.go:34-262) are all declared
    for n in ["gpu_init", "gpu_set_device", "gpu_maxthreads", "gpu_add", "gpu_madd", "gpu_madd2",
              "gpu_add_lin_anisotropy", "gpu_linear_combination", "gpu_linear_combination_many",
              "gpu_scale_dot_product", "gpu_add_constant", "gpu_reduce", "gpu_normalize_uniform",
              "gpu_normalize_map", "gpu_deltaM", "gpu_spintorque_deltaM", "gpu_add_local_fields",
              "gpu_add_local_fields_H_and_phi", "gpu_add_exch", "gpu_anal_fw_step", "gpu_kernelmul6",
              "gpu_kernelmul4", "gpu_copy_pad", "gpu_copy_unpad", "new_gpuFFT3dPlan_padded",
              "delete_gpuFFT3dPlan", "gpuFFT3dPlan_forward", "gpuFFT3dPlan_inverse", "gpu_gaussian_noise",
              "new_gpu_array", "free_gpu_array", "memcpy_gpu_dir", "gpu_stride_float", "gpu_override_stride",
              "gpu_zero", "gpu_usedmem", "gpu_print_properties_stdout"]:
        assert n in names, n


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libgpukern.so not built"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_tables_match_headers():
    import importlib

    gk = importlib.import_module("1_amd.gpukern")
    declared = declared_functions()
    bound = set(gk.GPUKERN) | set(gk.ENGINE) | set(gk.IO)
    assert declared == bound, (declared - bound, bound - declared)


def test_noalloc_forward_rejects_plans_beyond_its_scratch():
    """gpuFFT3dPlan_forward_noalloc checks its fixed scratch (4 Mi complex
    elements) before any device call: Z rows of half-length 3001 (prime) over
    2000 rows need 2000 x 3001 packed elements plus a Bluestein buffer, so it
    aborts with a message (the library's error behaviour).  Runs in a child
    process on the CPU; input and output are never touched."""
    import subprocess
    import sys

    code = ("import ctypes as C; lib = C.CDLL(%r); lib.new_gpuFFT3dPlan_padded.restype = C.c_void_p; "
            "p = lib.new_gpuFFT3dPlan_padded((C.c_int*3)(1, 2000, 3001), (C.c_int*3)(1, 2000, 6002)); "
            "lib.gpuFFT3dPlan_forward_noalloc(C.c_void_p(p), None, None)" % LIB)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "gpuFFT3dPlan_forward_noalloc" in r.stderr and "the maximum is 4194304" in r.stderr, r.stderr
