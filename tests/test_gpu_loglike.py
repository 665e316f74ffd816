"""Per-proposal log-likelihood of the HIP kernels against the reference's
golden vectors (tests/golden, from Group1/data/lba_data{2..6}.rda) and the
FP64 oracle (oracle/oracle.cpp)."""
import numpy as np
import pytest

import fixtures as fx
import oracle_lib as O

pytestmark = pytest.mark.gpu

# tolerance of the batched (FP32-transcendental) arithmetic, per proposal,
# relative to the FP64 value
MIXED_RTOL = 1e-6


def _golden_sets(fix):
    yield fix.sub_data, fix.sub_theta.T, fix.sub_ll
    for s in range(fix.n_subjects):
        yield fix.pop_data(s), fix.z["pop_theta"][s].T, fix.z["pop_ll"][s]


@pytest.mark.parametrize("name", fx.FIXTURES)
def test_fp64_matches_goldens(gpu, name):
    fix = fx.load(name)
    worst = 0.0
    n = 0
    for data, th, ll in _golden_sets(fix):
        got = gpu.loglike(fix.model, data, th, gpu.FP64_PLAIN)
        worst = max(worst, float(np.max(np.abs(got / ll - 1))))
        n += len(ll)
    print("%s: %d golden vectors, FP64 max rel %.2e" % (name, n, worst))
    assert worst <= 1e-12


@pytest.mark.parametrize("name", fx.FIXTURES)
def test_mixed_matches_goldens(gpu, name):
    fix = fx.load(name)
    errs = []
    for data, th, ll in _golden_sets(fix):
        got = gpu.loglike(fix.model, data, th, gpu.MIXED_BATCHED)
        errs.append(np.abs(got / ll - 1))
    e = np.concatenate(errs)
    print("%s: %d golden vectors, mixed max rel %.2e, %d above %.0e"
          % (name, len(e), e.max(), int((e > MIXED_RTOL).sum()), MIXED_RTOL))
    # every golden vector, prior draws with survivors at the FP64 rounding
    # floor included, within the per-proposal tolerance
    assert e.max() <= MIXED_RTOL


def test_mixed_near_posterior(gpu):
    fix = fx.load("lba_data6")
    worst = 0.0
    for s in range(0, fix.n_subjects, 4):
        data = fix.pop_data(s)
        th = fx.near_thetas(fix, 64, seed=100 + s)
        got = gpu.loglike(fix.model, data, th, gpu.MIXED_BATCHED)
        ref = np.array([O.lba_loglike(fix.omodel, data.rt, data.cell_off, t) for t in th])
        worst = max(worst, float(np.max(np.abs(got / ref - 1))))
    print("near-posterior mixed max rel %.2e" % worst)
    assert worst <= MIXED_RTOL


def test_fp64_near_posterior_matches_oracle(gpu):
    fix = fx.load("lba_data3")
    rng = np.random.default_rng(3)
    th = fix.sub_theta.T[:, :] * np.exp(rng.normal(0, 0.05, fix.sub_theta.T.shape))
    got = gpu.loglike(fix.model, fix.sub_data, th, gpu.FP64_PLAIN)
    ref = np.array([O.lba_loglike(fix.omodel, fix.sub_data.rt, fix.sub_data.cell_off, t) for t in th])
    assert np.allclose(got, ref, rtol=1e-12, atol=0)


def test_floored_and_empty_and_nan(gpu):
    fix = fx.load("lba_data6")
    th = fx.truth_vector(fix).copy()
    th[fix.model.pnames.index("t0")] = 50.0   # every trial has t <= 0
    n = len(fix.sub_data.rt)
    want = n * (np.log(1e-10) + np.log(1 - 1e-10))
    for prec in (gpu.FP64_PLAIN, gpu.MIXED_BATCHED):
        got = gpu.loglike(fix.model, fix.sub_data, th[None, :], prec)[0]
        assert abs(got - want) <= 1e-9 * abs(want)
    empty = gpu.SubjectData(np.zeros(0), np.zeros(fix.model.n_cell + 1, np.int64))
    for prec in (gpu.FP64_PLAIN, gpu.MIXED_BATCHED):
        assert gpu.loglike(fix.model, empty, fx.truth_vector(fix)[None, :], prec)[0] == 0.0
    bad = fx.truth_vector(fix).copy()
    bad[0] = np.nan
    for prec in (gpu.FP64_PLAIN, gpu.MIXED_BATCHED):
        assert np.isnan(gpu.loglike(fix.model, fix.sub_data, bad[None, :], prec)[0])
