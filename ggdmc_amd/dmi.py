"""Importers for the reference's R objects saved in `.rda` files (read with
ggdmc_amd.rdata): a `dmi` (data-model instance) as this engine's flattened
design and subject data, `prior` lists, and `posterior` objects as start
states for a restart.

The S4 slots read are those of ggdmcModel's dmi / model and ggdmc's
posterior (R/model-class.R:238-254): dmi@model, @data, @node_1_index;
model@model_boolean, @parameter_x_condition_names, @pnames, @constants,
@accumulators, @cell_names; posterior@theta, @summed_log_prior,
@log_likelihoods, @nmc, @thin, @nchain, @pnames, @start. The reference
builds its likelihood from the same slots (BuildDMI in
tests/testthat/Group1/data/6_lba_Bv_model.r:94-96; new_likelihood,
src/de2R.cpp:13).
"""
import numpy as np

from .rdata import RValue, load_rda
from .runs import LBAModel, Posterior, PriorSpec, SubjectData

# LBA parameter families, in the order of the flattened design's last axis.
FAMILIES = ("A", "B", "mean_v", "sd_v", "st0", "t0")


def _strs(v):
    return [str(x) for x in v.value]


def model_from_r(dmi):
    """dmi@model (+ dmi@node_1_index) -> LBAModel.

    model_boolean[c, k, j] marks the parameter condition k that feeds its
    family of accumulator j in cell c (exactly one per family); the family's
    value is the free parameter of that name (index in model@pnames) or the
    model constant of that name. The responding accumulator of cell c is
    node_1_index[c, 1] (the accumulator order is winner first)."""
    m = dmi.slot("model")
    pnames = _strs(m.slot("pnames"))
    pcond = _strs(m.slot("parameter_x_condition_names"))
    cst = m.slot("constants")
    consts = dict(zip(cst.names() or [], cst.value or []))
    accs = _strs(m.slot("accumulators"))
    cells = _strs(m.slot("cell_names"))
    n_cell, n_pc, n_acc = m.slot("model_boolean").dim()
    if n_pc != len(pcond) or n_acc != len(accs) or n_cell != len(cells):
        raise ValueError("model_boolean does not match the model's names")
    mb = np.array(m.slot("model_boolean").value, dtype=np.int32).reshape(
        (n_acc, n_pc, n_cell)).transpose(2, 1, 0)
    src = np.full((n_cell, n_acc, len(FAMILIES)), -2, dtype=np.int32)
    val = np.zeros((n_cell, n_acc, len(FAMILIES)))
    for c in range(n_cell):
        for j in range(n_acc):
            for k in np.nonzero(mb[c, :, j])[0]:
                name = pcond[k]
                fam = FAMILIES.index(name.split(".")[0])
                if src[c, j, fam] != -2:
                    raise ValueError("cell %s accumulator %s: two conditions for %s"
                                     % (cells[c], accs[j], FAMILIES[fam]))
                if name in pnames:
                    src[c, j, fam] = pnames.index(name)
                else:
                    if name not in consts:
                        raise ValueError("%s is neither a free parameter nor a constant" % name)
                    src[c, j, fam] = -1
                    val[c, j, fam] = consts[name]
    if (src == -2).any():
        raise ValueError("a parameter family without a condition in model_boolean")
    n1 = np.array(dmi.slot("node_1_index").value, dtype=np.int32).reshape((n_acc, n_cell)).T
    # each row is the accumulator order of a cell (a permutation): its
    # smallest entry gives the index base
    winner = (n1[:, 0] - n1.min()).astype(np.int32)
    for c in range(n_cell):
        if cells[c].split(".")[-1] != accs[winner[c]]:
            raise ValueError("cell %s: node_1_index names accumulator %s as the response"
                             % (cells[c], accs[winner[c]]))
    return LBAModel(winner, src, val, pnames, cells)


def data_from_r(dmi, cells):
    """dmi@data (a named list of per-cell response times; cells without
    trials are absent) -> SubjectData in model@cell_names order."""
    d = dmi.slot("data")
    by_name = dict(zip(d.names() or [], d.value or []))
    return SubjectData.from_cells([np.array(by_name[c].value if c in by_name else [], float)
                                   for c in cells])


def dmi_from_r(dmi):
    """A dmi S4 object -> (LBAModel, SubjectData), the pair the sampling API
    takes for a subject."""
    model = model_from_r(dmi)
    return model, data_from_r(dmi, model.cell_names)


def prior_from_r(plist, pnames=None):
    """A p_prior / h_prior named list (entries p0, p1, lower, upper, dist_id)
    -> PriorSpec, in pnames order (default: the list's order). An NA bound is
    unbounded."""
    names = plist.names()
    order = list(pnames) if pnames is not None else names
    cols = {k: [] for k in ("p0", "p1", "lower", "upper", "dist_id")}
    for p in order:
        e = plist[names.index(str(p))]
        for k in cols:
            cols[k].append(e[k].value[0])
    lo = np.array(cols["lower"], float)
    up = np.array(cols["upper"], float)
    lo[np.isnan(lo)] = -np.inf
    up[np.isnan(up)] = np.inf
    return PriorSpec(np.array(cols["dist_id"], np.int32), np.array(cols["p0"], float),
                     np.array(cols["p1"], float), lo, up, [str(p) for p in order])


def priors_from_r(prior):
    """A prior S4 object (ggdmcPrior::set_priors: @p_prior, @h_prior,
    @nparameter, @pnames) -> the sampling API's priors dict."""
    pn = _strs(prior.slot("pnames")) if "pnames" in prior.attrs else None
    hp = prior.attrs.get("h_prior")
    has_h = isinstance(hp, RValue) and hp.kind == "list" and len(hp.value) > 0
    p_names = prior.slot("p_prior").names()
    out = {"p_prior": prior_from_r(prior.slot("p_prior"), p_names),
           "h_prior": prior_from_r(hp, pn) if has_h else None}
    out["nparameter"] = int(prior.slot("nparameter").value[0])
    return out


def posterior_from_r(post):
    """A posterior S4 object -> Posterior (theta [npar, nchain, nmc],
    summed_log_prior / log_likelihoods [nchain, nmc])."""
    P, n, nmc = post.slot("theta").dim()
    th = np.array(post.slot("theta").value, float).reshape((nmc, n, P)).transpose(2, 1, 0)
    lp = np.array(post.slot("summed_log_prior").value, float).reshape((nmc, n)).T
    ll = np.array(post.slot("log_likelihoods").value, float).reshape((nmc, n)).T

    def one(name, default):
        v = post.attrs.get(name)
        return int(v.value[0]) if v is not None else default

    return Posterior(th, lp, ll, one("nmc", nmc), one("thin", 1), one("nchain", n),
                     _strs(post.slot("pnames")) if "pnames" in post.attrs else [],
                     one("start", 1))


def first_state(post):
    """The start state a run takes from a posterior built by
    initialise_theta / initialise_phi: column 1 (theta [P, nchain], lp, ll)."""
    p = posterior_from_r(post) if isinstance(post, RValue) else post
    return p.theta[:, :, 0].copy(), p.summed_log_prior[:, 0].copy(), p.log_likelihoods[:, 0].copy()


def load(path):
    """Every object of an `.rda` file, decoded ({name: RValue})."""
    return load_rda(path)
