"""ctypes binding of tests/native/libshim_abi.so (the shim's ABI path in C;
test infrastructure)."""
import ctypes as C
import os

import numpy as np

from ggdmc_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "native", "libshim_abi.so")
_d = C.POINTER(C.c_double)


def load():
    L = C.CDLL(PATH)
    L.shim_start_column.restype = C.c_int
    L.shim_start_column.argtypes = [_d, C.c_int, C.c_int, C.c_int]
    L.shim_run.restype = C.c_int
    L.shim_run.argtypes = [C.POINTER(_lib.Config), C.POINTER(_lib.Model), C.POINTER(_lib.Prior),
                           C.POINTER(_lib.Prior), C.c_int32, C.c_int32, C.POINTER(_lib.Subject),
                           _d, C.c_int32, _d, _d, _d, C.POINTER(_d), C.POINTER(_d),
                           C.POINTER(_d), C.c_int32, C.c_uint64, C.c_int32, C.c_int32,
                           C.POINTER(_lib.Samples), C.POINTER(_lib.Samples)]
    return L


def r_samples(theta, lp, ll, nmc_in, filled=1):
    """A samples object as R holds it: theta [P x n x nmc_in], lp / ll
    [n x nmc_in] (column-major), the first `filled` columns from the given
    states (arrays [.., filled]), the rest as set_up_new_samples leaves them
    (NaN parameters, -Inf densities)."""
    P, n = theta.shape[:2]
    th = np.full((P, n, nmc_in), np.nan, order="F")
    a = np.full((n, nmc_in), -np.inf, order="F")
    b = np.full((n, nmc_in), -np.inf, order="F")
    th[:, :, :filled] = theta.reshape(P, n, -1)[:, :, :filled]
    a[:, :filled] = lp.reshape(n, -1)[:, :filled]
    b[:, :filled] = ll.reshape(n, -1)[:, :filled]
    return th, a, b


def r_out(P, n, nmc, keep):
    th = np.zeros((P, n, nmc), order="F")
    lp = np.zeros((n, nmc), order="F")
    ll = np.zeros((n, nmc), order="F")
    keep.extend([th, lp, ll])
    return (th, lp, ll), _lib.Samples(_lib.ptr(th), _lib.ptr(lp), _lib.ptr(ll))
