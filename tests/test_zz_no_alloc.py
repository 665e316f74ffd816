"""ggdmc_loglike_noalloc (the allocation-free variant of ggdmc_loglike, staged
in the library's fixed device array) gives the bits of ggdmc_loglike, on two
calls in a row whose second is smaller (the array keeps the first call's
contents), and rejects inputs larger than the array."""
import ctypes as C

import numpy as np
import pytest

import fixtures as fx
from ggdmc_amd import _lib
from ggdmc_amd.runs import SubjectData

pytestmark = pytest.mark.gpu


def _call(fn, model, data, thetas, precision):
    k = _lib.Keep()
    cm = model.c_struct(k)
    cs = data.c_struct(k)
    th = k.arr(np.atleast_2d(thetas), np.float64)
    out = np.full(th.shape[0], np.nan)
    _lib.check(fn(C.byref(cm), C.byref(cs), _lib.ptr(th), th.shape[0], int(precision),
                  _lib.ptr(out)))
    return out


def test_noalloc_equals_loglike_on_two_calls(gpu):
    L = _lib.lib()
    big = fx.load("lba_data6")          # 768 trials, 2 accumulators, 13 parameters
    small = fx.load("lba_data2")        # 256 trials, 2 accumulators, 8 parameters
    cases = [(big, big.sub_data, big.sub_theta[:, :32].T),
             (small, small.pop_data(3), small.z["pop_theta"][3][:, :5].T)]
    for prec in (gpu.FP64_PLAIN, gpu.MIXED_BATCHED):
        for fix, data, th in cases:
            print("noalloc: %s, %d trials, %d thetas, precision %d"
                  % (fix.name, len(data.rt), th.shape[0], prec), flush=True)
            want = _call(L.ggdmc_loglike, fix.model, data, th, prec)
            got = _call(L.ggdmc_loglike_noalloc, fix.model, data, th, prec)
            assert np.all(np.isfinite(want))
            assert np.array_equal(got, want), (got, want)


def test_noalloc_rejects_inputs_beyond_the_array(gpu):
    fix = fx.load("lba_data6")
    nc = fix.model.n_cell
    n = 2_000_000                        # 20 MB of trials > the 16 MiB array
    off = np.linspace(0, n, nc + 1).astype(np.int64)
    data = SubjectData(np.full(n, 0.6), off)
    with pytest.raises(gpu.GgdmcError, match="staging array"):
        _call(_lib.lib().ggdmc_loglike_noalloc, fix.model, data, fix.sub_theta[:, :1].T,
              gpu.FP64_PLAIN)
    # the same input still runs through the allocating entry point's checks,
    # and the array is usable afterwards
    th = fix.sub_theta[:, :4].T
    assert np.array_equal(gpu.loglike(fix.model, fix.sub_data, th, gpu.FP64_PLAIN, noalloc=True),
                          gpu.loglike(fix.model, fix.sub_data, th, gpu.FP64_PLAIN))
