"""Golden fixtures (tests/golden/lba_data{2..6}.npz, written by
tests/golden/extract_goldens.py from the reference's Group1/data/*.rda) as
product-side and oracle-side objects, plus synthetic inputs."""
import os
import sys
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle_lib as O  # noqa: E402
from ggdmc_amd.runs import LBAModel, PriorSpec, SubjectData  # noqa: E402

GOLDEN = os.path.join(HERE, "golden")
FIXTURES = ["lba_data2", "lba_data3", "lba_data4", "lba_data5", "lba_data6"]

# generating values of the B x v design (Group1/data/6_lba_Bv_model.r:41-85)
BV_TRUTH = {"A": 0.25, "B.left.red": 2.33, "B.right.red": 2.13, "B.left.blue": 2.55,
            "B.right.blue": 2.35, "t0": 0.05, "mean_v.high.true": 2.1,
            "mean_v.moderate.true": 3.1, "mean_v.low.true": 4.1, "mean_v.high.false": 1.82,
            "mean_v.moderate.false": 1.70, "mean_v.low.false": 1.5, "sd_v.true": 0.2}
BV_SCALE = {"A": 0.2, "B.left.red": .23, "B.right.red": .23, "B.left.blue": .255,
            "B.right.blue": .235, "t0": 0.05, "mean_v.high.true": .25, "mean_v.moderate.true": .35,
            "mean_v.low.true": .45, "mean_v.high.false": .182, "mean_v.moderate.false": .170,
            "mean_v.low.false": .15, "sd_v.true": 0.2}


def _prior(z, key, names):
    return PriorSpec(z[key + "_dist_id"].astype(np.int32), z[key + "_p0"], z[key + "_p1"],
                     z[key + "_lower"], z[key + "_upper"], list(names))


def oprior(p):
    return O.Prior(p.dist, p.p0, p.p1, p.lower, p.upper)


@dataclass
class Fixture:
    name: str
    z: object
    model: LBAModel
    omodel: object
    P: int
    sub_data: SubjectData
    sub_theta: np.ndarray
    sub_ll: np.ndarray
    sub_lp: np.ndarray
    sub_prior: PriorSpec
    oprior: object
    pop_prior: PriorSpec
    hyper_prior: PriorSpec

    @property
    def n_subjects(self):
        return self.z["pop_theta"].shape[0]

    def pop_data(self, s):
        so = self.z["pop_subject_off"]
        return SubjectData(self.z["pop_rt"][so[s]:so[s + 1]], self.z["pop_cell_off"][s])


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    pn = [str(x) for x in z["design_pnames"]]
    model = LBAModel(z["design_winner"].astype(np.int32), z["design_par_src"].astype(np.int32),
                     z["design_par_const"], pn, [str(x) for x in z["design_cell_names"]])
    om = O.Model(model.winner, model.par_src, model.par_const, len(pn))
    sp = _prior(z, "sub_prior", pn)
    return Fixture(name, z, model, om, len(pn), SubjectData(z["sub_rt"], z["sub_cell_off"]),
                   z["sub_theta"], z["sub_ll"], z["sub_lp"], sp, oprior(sp),
                   _prior(z, "pop_prior", pn),
                   _prior(z, "hyper_prior", [str(x) for x in z["phi_pnames"]]))


def truth_vector(fix):
    return np.array([BV_TRUTH[p] for p in fix.model.pnames])


def near_start(fix, nchain, seed, scale=0.1, data=None, prior=None):
    """Start state near the generating values of the B x v design: theta
    [P, nchain] with lp / ll from the oracle (FP64)."""
    rng = np.random.default_rng(seed)
    tv = truth_vector(fix)
    data = data or fix.sub_data
    pr = oprior(prior or fix.sub_prior)
    th = tv[:, None] * np.exp(rng.normal(0, scale, (fix.P, nchain)))
    lp = np.array([O.sumlogprior(pr, th[:, k]) for k in range(nchain)])
    ll = np.array([O.lba_loglike(fix.omodel, data.rt, data.cell_off, th[:, k])
                   for k in range(nchain)])
    return th, lp, ll


def near_thetas(fix, n, seed, scale=0.15):
    rng = np.random.default_rng(seed)
    tv = truth_vector(fix)
    return tv[None, :] * np.exp(rng.normal(0, scale, (n, fix.P)))


def hier_start(fix, S, nchain, seed, scale=0.1):
    """Near-posterior start of a hierarchical run on the first S subjects:
    phi [2P, n] (loc, scale), subject states [(theta [P, n], lp, ll)]; lp / ll
    from the oracle (subject lp is tnorm under phi chain c)."""
    rng = np.random.default_rng(seed)
    tv = truth_vector(fix)
    sc = np.array([BV_SCALE[p] for p in fix.model.pnames])
    P = fix.P
    phi = np.concatenate([tv[:, None] * np.exp(rng.normal(0, scale, (P, nchain))),
                          sc[:, None] * np.exp(rng.normal(0, scale, (P, nchain)))])
    hp = oprior(fix.hyper_prior)
    pp = fix.pop_prior
    subs = []
    for s in range(S):
        d = fix.pop_data(s)
        th = tv[:, None] * np.exp(rng.normal(0, scale, (P, nchain)))
        lp = np.empty(nchain)
        ll = np.empty(nchain)
        for c in range(nchain):
            pc = O.Prior(np.full(P, 1, np.int32), phi[:P, c], phi[P:, c], pp.lower, pp.upper)
            lp[c] = O.sumlogprior(pc, th[:, c])
            ll[c] = O.lba_loglike(fix.omodel, d.rt, d.cell_off, th[:, c])
        subs.append((th, lp, ll))
    plp = np.array([O.sumlogprior(hp, phi[:, c]) for c in range(nchain)])
    opp = oprior(pp)
    pll = np.array([O.sumloghlike(opp, phi[:, c], np.stack([s[0][:, c] for s in subs]))
                    if subs else 0.0 for c in range(nchain)])
    return (phi, plp, pll), subs
