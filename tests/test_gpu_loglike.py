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


def test_mixed_far_tail_winner_trials(gpu):
    # fast responses of a winner far from its threshold (z2 = 5 .. 7), the
    # density near the 1e-10 floor: there the reference's v * (Phi(z1) -
    # Phi(z2)) is quantised by its own rounding near 1, which the batched path
    # reproduces or evaluates with the reference formula. One trial of cell
    # s1.r1 (r1 wins) under 4096 parameter vectors, each term against the oracle
    from ggdmc_amd.model import build_model
    m = build_model(p_map={"A": ["1"], "B": ["1"], "t0": ["1"], "mean_v": ["M"], "sd_v": ["M"],
                           "st0": ["1"]}, match_map={"M": {"s1": "r1", "s2": "r2"}},
                    factors={"S": ["s1", "s2"]}, constants={"st0": 0.0, "sd_v.false": 0.3},
                    accumulators=["r1", "r2"])
    assert m.pnames == ["A", "B", "mean_v.false", "mean_v.true", "sd_v.true", "t0"]
    data = gpu.SubjectData.from_cells([[0.35], [], [], []])
    rng = np.random.default_rng(8)
    n = 4096
    A, v, s = rng.uniform(0.2, 1.0, n), rng.uniform(3, 12, n), rng.uniform(0.5, 1.5, n)
    t = 0.35 - 0.2
    B = t * (v + rng.uniform(5.0, 7.0, n) * s)          # z2 = B / (t s) - v / s
    th = np.stack([A, B, np.full(n, 1.0), v, s, np.full(n, 0.2)], axis=1)
    got = gpu.loglike(m, data, th, gpu.MIXED_BATCHED)
    om = O.Model(m.winner, m.par_src, m.par_const, m.npar)
    ref = np.array([O.lba_loglike(om, data.rt, data.cell_off, x) for x in th])
    far = (ref > np.log(2e-10)) & (ref < np.log(1e-6))
    d = np.abs(got - ref)
    # where that rounding could move the winner density by more than 1e-5
    # (|v| 2^-52 > 1e-5 |numerator|), the reference's own value is reproduced
    from scipy.special import erfc
    b = A + B
    z1, z2 = (b - t * v) / (t * s), (b - t * v - A) / (t * s)
    Q1, Q2 = 0.5 * erfc(z1 * np.sqrt(0.5)), 0.5 * erfc(z2 * np.sqrt(0.5))
    num = v * ((1.0 - Q1) - (1.0 - Q2)) + s * (np.exp(-0.5 * z2 * z2) - np.exp(-0.5 * z1 * z1)) / np.sqrt(2 * np.pi)
    coarse = far & (2.0 ** -52 * v > 1.2e-5 * np.abs(num))
    print("far-tail terms: %d, max |dLL| %.2e; %d where the rounding matters, max %.2e"
          % (far.sum(), d[far].max(), coarse.sum(), d[coarse].max()))
    assert far.sum() > 2000 and coarse.sum() > 50
    assert d.max() <= 1e-5 and d[coarse].max() <= 1e-7
