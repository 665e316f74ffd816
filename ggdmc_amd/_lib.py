"""ctypes binding of libggdmc_hip.so (include/ggdmc_hip.h).

The library is the only compute path: importing a function that needs it
raises if ggdmc_amd/libggdmc_hip.so is absent, and every call that would run
on the GPU returns GGDMC_ENODEV without one. There is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libggdmc_hip.so")

GGDMC_OK = 0
GGDMC_EINVAL = -1
GGDMC_EHIP = -2
GGDMC_ENOMEM = -3
GGDMC_ENODEV = -4

FP64_PLAIN = 0
MIXED_BATCHED = 1
FP64_BATCHED = 2

KIND_SUBJECT = 0
KIND_HYPER = 1
KIND_HIER = 2

DIST_TNORM = 1
DIST_UNIF = 6

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)


class Config(C.Structure):
    _fields_ = [("nmc", C.c_int32), ("thin", C.c_int32), ("nchain", C.c_int32),
                ("pop_migration_prob", C.c_double), ("sub_migration_prob", C.c_double),
                ("gamma_precursor", C.c_double), ("rp", C.c_double),
                ("is_hblocked", C.c_int32), ("is_pblocked", C.c_int32),
                ("precision", C.c_int32), ("sweep_groups", C.c_int32),
                ("trials_per_tile", C.c_int32), ("chains_per_item", C.c_int32),
                ("launch_mode", C.c_int32), ("iter_offset", C.c_uint64),
                ("pop_debug", C.c_int32), ("sub_debug", C.c_int32),
                ("waves_per_simd", C.c_int32)]


class TraceRecord(C.Structure):
    _fields_ = [("job", C.c_int32), ("subject", C.c_int32), ("iteration", C.c_uint32),
                ("sweep", C.c_int32), ("kind", C.c_int32), ("step", C.c_int32),
                ("chain", C.c_int32), ("target", C.c_int32), ("tmp_lp", C.c_double),
                ("tmp_ll", C.c_double), ("cur_post", C.c_double), ("accepted", C.c_int32),
                ("reserved", C.c_int32)]


class Model(C.Structure):
    _fields_ = [("n_acc", C.c_int32), ("n_cell", C.c_int32), ("npar", C.c_int32),
                ("winner", _i32), ("par_src", _i32), ("par_const", _d)]


class Prior(C.Structure):
    _fields_ = [("npar", C.c_int32), ("dist", _i32), ("p0", _d), ("p1", _d),
                ("lower", _d), ("upper", _d)]


class Subject(C.Structure):
    _fields_ = [("ntrial", C.c_int64), ("rt", _d), ("cell_off", _i64)]


class State(C.Structure):
    _fields_ = [("theta", _d), ("lp", _d), ("ll", _d)]


class Samples(C.Structure):
    _fields_ = [("theta", _d), ("lp", _d), ("ll", _d)]


class Job(C.Structure):
    _fields_ = [("kind", C.c_int32), ("seed", C.c_uint64), ("nparameter_de", C.c_int32),
                ("p_prior", C.POINTER(Prior)), ("h_prior", C.POINTER(Prior)),
                ("nsubject", C.c_int32), ("subjects", C.POINTER(Subject)),
                ("hyper_data", _d), ("phi_start", C.POINTER(State)),
                ("subject_starts", C.POINTER(State)), ("phi_out", C.POINTER(Samples)),
                ("subject_outs", C.POINTER(Samples))]


class GgdmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("ggdmc error %d: %s" % (code, msg))
        self.code = code


_LIB = None


def lib():
    """Loads libggdmc_hip.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("ggdmc_amd/libggdmc_hip.so is not built: run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    L.ggdmc_last_error.restype = C.c_char_p
    L.ggdmc_device_count.restype = C.c_int
    L.ggdmc_run_batch.argtypes = [C.POINTER(Config), C.POINTER(Model), C.POINTER(Job), C.c_int32]
    L.ggdmc_run_subject.argtypes = [C.POINTER(Config), C.POINTER(Model), C.POINTER(Prior),
                                    C.POINTER(Subject), C.POINTER(State), C.c_uint64,
                                    C.POINTER(Samples)]
    L.ggdmc_run_hyper.argtypes = [C.POINTER(Config), C.POINTER(Prior), C.POINTER(Prior),
                                  C.c_int32, _d, C.POINTER(State), C.c_uint64,
                                  C.POINTER(Samples)]
    L.ggdmc_run.argtypes = [C.POINTER(Config), C.POINTER(Model), C.POINTER(Prior),
                            C.POINTER(Prior), C.c_int32, C.POINTER(Subject), C.POINTER(State),
                            C.POINTER(State), C.c_uint64, C.POINTER(Samples),
                            C.POINTER(Samples)]
    L.ggdmc_loglike.argtypes = [C.POINTER(Model), C.POINTER(Subject), _d, C.c_int32, C.c_int32,
                                _d]
    L.ggdmc_loglike_noalloc.argtypes = L.ggdmc_loglike.argtypes
    L.ggdmc_session_create.argtypes = [C.POINTER(Config), C.POINTER(Model), C.POINTER(Job),
                                       C.c_int32, C.POINTER(C.c_void_p)]
    L.ggdmc_session_advance.argtypes = [C.c_void_p, C.c_int32]
    L.ggdmc_session_elapsed_ms.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
    L.ggdmc_session_download.argtypes = [C.c_void_p]
    L.ggdmc_session_profile.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double),
                                        C.POINTER(C.c_int32)]
    L.ggdmc_session_trial_evals.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.ggdmc_session_decomposition.argtypes = [C.c_void_p, _i32, _i32, _i32, _i32]
    L.ggdmc_session_waves_per_simd.argtypes = [C.c_void_p, _i32]
    L.ggdmc_session_trace.argtypes = [C.c_void_p, C.POINTER(C.POINTER(TraceRecord)),
                                      C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.ggdmc_last_trace.argtypes = [C.POINTER(C.POINTER(TraceRecord)), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64)]
    L.ggdmc_session_destroy.argtypes = [C.c_void_p]
    L.ggdmc_session_destroy.restype = None
    _LIB = L
    return L


TRACE_DTYPE = np.dtype([("job", np.int32), ("subject", np.int32), ("iteration", np.uint32),
                        ("sweep", np.int32), ("kind", np.int32), ("step", np.int32),
                        ("chain", np.int32), ("target", np.int32), ("tmp_lp", np.float64),
                        ("tmp_ll", np.float64), ("cur_post", np.float64),
                        ("accepted", np.int32), ("reserved", np.int32)])


def trace_array(ptr_, n):
    """A copy of n trace records at ptr_ as a numpy structured array."""
    if n <= 0:
        return np.zeros(0, TRACE_DTYPE)
    buf = C.string_at(C.cast(ptr_, C.c_void_p), int(n) * C.sizeof(TraceRecord))
    return np.frombuffer(buf, TRACE_DTYPE).copy()


def check(rc):
    if rc != GGDMC_OK:
        raise GgdmcError(rc, lib().ggdmc_last_error().decode())


def ptr(a, t=_d):
    return a.ctypes.data_as(t)


def device_count():
    return lib().ggdmc_device_count()


class Keep:
    """Holds numpy arrays alive while C structs point into them."""

    def __init__(self):
        self.items = []

    def arr(self, a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        self.items.append(a)
        return a
