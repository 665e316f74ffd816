"""The product's importer of the reference's R objects (ggdmc_amd.dmi over
ggdmc_amd.rdata) against the derived golden fixtures (tests/golden, written
from the same .rda files): the flattened design, the subjects' trials, the
priors and the start states. Needs the reference's .rda files, so it runs
where they are mounted and is skipped elsewhere."""
import os

import numpy as np
import pytest

import fixtures as fx
from ggdmc_amd import dmi as D

REF = "/root/reference/tests/testthat/Group1/data"
pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason="reference .rda files not mounted")


@pytest.mark.parametrize("name", fx.FIXTURES)
def test_import_matches_the_fixtures(name):
    fix = fx.load(name)
    d = D.load(os.path.join(REF, name + ".rda"))
    model, data = D.dmi_from_r(d["sub_dmis"][0])
    assert model.pnames == fix.model.pnames and model.cell_names == fix.model.cell_names
    assert np.array_equal(model.winner, fix.model.winner)
    assert np.array_equal(model.par_src, fix.model.par_src)
    assert np.array_equal(model.par_const, fix.model.par_const)
    assert np.array_equal(data.rt, fix.sub_data.rt)
    assert np.array_equal(data.cell_off, fix.sub_data.cell_off)
    th, lp, ll = D.first_state(d["sub_samples"])
    assert np.array_equal(th, fix.sub_theta) and np.array_equal(ll, fix.sub_ll)
    assert np.array_equal(lp, fix.sub_lp)
    sp = D.priors_from_r(d["sub_priors"])["p_prior"]
    assert np.array_equal(sp.p0, fix.sub_prior.p0) and np.array_equal(sp.dist, fix.sub_prior.dist)
    pop = D.priors_from_r(d["pop_priors"])
    assert pop["nparameter"] == 2 * fix.P
    assert np.array_equal(pop["h_prior"].p1, fix.hyper_prior.p1)
    assert np.array_equal(pop["p_prior"].lower, fix.pop_prior.lower)
    dmis = [D.dmi_from_r(x) for x in d["pop_dmis"].value]
    for s in (0, len(dmis) - 1):
        want = fix.pop_data(s)
        assert np.array_equal(dmis[s][1].rt, want.rt)
        assert np.array_equal(dmis[s][1].cell_off, want.cell_off)
    phi = D.posterior_from_r(d["pop_samples"]["phi"])
    assert np.array_equal(phi.theta[:, :, 0], fix.z["phi_theta"])
    assert phi.nchain == fix.z["phi_theta"].shape[1]
