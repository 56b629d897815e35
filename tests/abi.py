"""ctypes binding of include/mars_hip.h used by the tests (the C ABI is the product boundary)."""

import ctypes as C
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mars_hip.h")
LIB = os.path.join(ROOT, "mars_amd", "libmars_hip.so")


class Best(C.Structure):
    _fields_ = [("sse1", C.c_double), ("xcol1", C.c_int), ("bcol1", C.c_int), ("sse2", C.c_double),
                ("xcol2", C.c_int), ("bcol2", C.c_int), ("knot2", C.c_int), ("cut2", C.c_double),
                ("sse0", C.c_double)]


def declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(mars_hip_\w+)\s*\(", txt)))


def load():
    lib = C.CDLL(LIB)
    P, sz, d, i = C.c_void_p, C.c_size_t, C.c_double, C.c_int
    fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    lib.mars_hip_last_error.restype = C.c_char_p
    lib.mars_hip_context_create.argtypes = [sz, sz, d, C.POINTER(P)]
    lib.mars_hip_context_destroy.argtypes = [P]
    lib.mars_hip_context_batch_capacity.argtypes = [P, C.POINTER(sz)]
    lib.mars_hip_context_sync_basis.argtypes = [P, sz, P, sz, P, sz]
    lib.mars_hip_context_set_target.argtypes = [P, P]
    lib.mars_hip_context_sync_xcols.argtypes = [P, sz, P, sz, P, P, sz]
    lib.mars_hip_orthonormalize.argtypes = [P, sz, sz, P, P, P, sz, P, C.POINTER(C.c_long)]
    lib.mars_hip_orthonormalize_batch.argtypes = [P, sz, sz, P, P, P, C.POINTER(C.c_long)]
    lib.mars_hip_fit_create.argtypes = [P, sz, P, P, sz, sz, sz, i, C.POINTER(P)]
    lib.mars_hip_fit_info.argtypes = [P, C.POINTER(i), C.POINTER(d), C.POINTER(d)]
    lib.mars_hip_fit_set_form.argtypes = [P, i]
    lib.mars_hip_fit_eval.argtypes = [P, P, sz, sz, i, i, i, P, P, P, C.POINTER(Best)]
    lib.mars_hip_fit_append.argtypes = [P, C.c_char, i, i, C.c_float, C.POINTER(d)]
    lib.mars_hip_fit_basis.argtypes = [P, P, P, C.POINTER(sz)]
    lib.mars_hip_stats.argtypes = [P, C.POINTER(C.c_long), C.POINTER(d)]
    lib.mars_hip_expand.argtypes = [P, sz, sz, sz, sz, C.c_char_p, P, P, P, P, sz]
    lib.mars_hip_expand_f64.argtypes = [P, sz, sz, sz, sz, C.c_char_p, P, P, P, P, sz]
    lib.mars_hip_gram.argtypes = [P, sz, sz, sz, P, sz, sz, P, sz]
    lib.mars_hip_gram_f64.argtypes = [P, sz, sz, sz, P, sz, sz, P, sz]
    del fp
    return lib


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def check(lib, rc):
    if rc != 0:
        raise RuntimeError("mars_hip error %d: %s" % (rc, lib.mars_hip_last_error().decode()))
