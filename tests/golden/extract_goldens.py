"""Write the likelihood/prior golden vectors from the reference's own fixtures.

Reads tests/testthat/Group1/data/lba_data{2..6}.rda of the reference (the
initial states written by initialise_theta / initialise_phi,
R/phi.R:141-204 and :256-332; generators Group1/data/{2..6}_lba_*.r) and saves,
per fixture, an .npz holding only data: the flattened design, the trial RTs,
the prior specifications and the (theta, log-likelihood, log-prior) triples
that the reference computed. Run where /root/reference is mounted:

    python tests/golden/extract_goldens.py

lba_data0/1 are skipped: their non-floored log-likelihoods come from an older
likelihood and do not match the fixtures' own data.
"""
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from ggdmc_amd.rdata import load_rda  # noqa: E402

REF_DATA = "/root/reference/tests/testthat/Group1/data"
OUT_DIR = os.path.dirname(os.path.abspath(__file__))

# Parameter families of an LBA accumulator, in the order the flattened design
# stores them (one source per family per (cell, accumulator)).
FAMILIES = ("A", "B", "mean_v", "sd_v", "st0", "t0")


def flatten_design(dmi):
    """dmi@model -> (winner[n_cell], par_src[n_cell, n_acc, 6], par_const[...]).

    model_boolean[c, k, j] marks which parameter-condition k feeds family
    family(k) of accumulator j in cell c; exactly one per family.
    par_src holds the index into the free-parameter vector (pnames order) or
    -1 when the value is the model constant stored in par_const.
    """
    m = dmi.slot("model")
    pnames = list(m.slot("pnames").value)
    pcond = list(m.slot("parameter_x_condition_names").value)
    cnames = m.slot("constants").names() or []
    cvals = m.slot("constants").value if cnames else []
    consts = dict(zip(cnames, cvals))
    accs = list(m.slot("accumulators").value)
    cells = list(m.slot("cell_names").value)
    n_cell, n_pc, n_acc = m.slot("model_boolean").dim()
    assert n_pc == len(pcond) and n_acc == len(accs) and n_cell == len(cells)
    mb = np.array(m.slot("model_boolean").value, dtype=np.int32).reshape(
        (n_acc, n_pc, n_cell)).transpose(2, 1, 0)
    par_src = np.full((n_cell, n_acc, len(FAMILIES)), -2, dtype=np.int32)
    par_const = np.zeros((n_cell, n_acc, len(FAMILIES)))
    for c in range(n_cell):
        for j in range(n_acc):
            for k in np.nonzero(mb[c, :, j])[0]:
                name = pcond[k]
                fam = FAMILIES.index(name.split(".")[0])
                assert par_src[c, j, fam] == -2, "two conditions for one family"
                if name in pnames:
                    par_src[c, j, fam] = pnames.index(name)
                else:
                    par_src[c, j, fam] = -1
                    par_const[c, j, fam] = consts[name]
    assert (par_src != -2).all(), "a family without a parameter condition"
    n1 = np.array(dmi.slot("node_1_index").value, dtype=np.int32).reshape(
        (n_acc, n_cell)).T
    winner = n1[:, 0].copy()
    for c in range(n_cell):
        assert cells[c].split(".")[-1] == accs[winner[c]]
    spec = {
        "p_map": {k: list(v.value) for k, v in zip(m.slot("parameter_map").names(),
                                                   m.slot("parameter_map").value)},
        "factors": {k: list(v.value) for k, v in zip(m.slot("factors").names(),
                                                     m.slot("factors").value)},
        "match_map": {k: {kk: vv.value[0] for kk, vv in zip(v.names(), v.value)}
                      for k, v in zip(m.slot("match_map").names(), m.slot("match_map").value)},
        "constants": dict(zip(cnames, cvals)),
        "accumulators": accs,
        "parameter_x_condition_names": pcond,
    }
    return dict(winner=winner, par_src=par_src, par_const=par_const,
                spec_json=np.array(json.dumps(spec, ensure_ascii=False)),
                n_acc=np.int32(n_acc), n_cell=np.int32(n_cell),
                pnames=np.array(pnames), cell_names=np.array(cells),
                accumulators=np.array(accs))


def subject_trials(dmi, cells):
    """Cell-sorted RTs and cell offsets [n_cell + 1] (empty cells are absent
    from dmi@data)."""
    data = dmi.slot("data")
    by_name = dict(zip(data.names(), data.value))
    rts, offs = [], [0]
    for c in cells:
        v = by_name.get(str(c))
        x = np.array(v.value if v is not None else [], dtype=np.float64)
        rts.append(x)
        offs.append(offs[-1] + len(x))
    return np.concatenate(rts), np.array(offs, dtype=np.int64)


def prior_spec(prior_list, pnames):
    """p_prior / h_prior named list -> arrays in pnames order."""
    names = prior_list.names()
    out = {k: [] for k in ("p0", "p1", "lower", "upper", "dist_id", "log_p")}
    for p in pnames:
        e = prior_list[names.index(str(p))]
        for k in out:
            out[k].append(e[k].value[0])
    return {k: np.array(v, dtype=np.float64) for k, v in out.items()}


def first_state(post):
    """posterior S4 -> theta[:, :, 0] [P, nchain], ll[:, 0], lp[:, 0]."""
    P, nchain, nmc = post.slot("theta").dim()
    th = np.array(post.slot("theta").value[:P * nchain]).reshape((nchain, P)).T
    ll = np.array(post.slot("log_likelihoods").value[:nchain])
    lp = np.array(post.slot("summed_log_prior").value[:nchain])
    return th, ll, lp


def extract(idx):
    d = load_rda(os.path.join(REF_DATA, "lba_data%d.rda" % idx))
    sub_dmi = d["sub_dmis"][0]
    design = flatten_design(sub_dmi)
    cells = design["cell_names"]
    pnames = design["pnames"]
    out = {"design_" + k: v for k, v in design.items()}

    rt, off = subject_trials(sub_dmi, cells)
    th, ll, lp = first_state(d["sub_samples"])
    out.update(sub_rt=rt, sub_cell_off=off, sub_theta=th, sub_ll=ll, sub_lp=lp)
    for k, v in prior_spec(d["sub_priors"].slot("p_prior"), pnames).items():
        out["sub_prior_" + k] = v

    pop = d["pop_samples"]
    subs = pop["subject_theta"].value
    dmis = d["pop_dmis"].value
    S = len(subs)
    rts, offs, ths, lls, lps = [], [], [], [], []
    for s in range(S):
        assert list(flatten_design(dmis[s])["pnames"]) == list(pnames)
        r, o = subject_trials(dmis[s], cells)
        t, a, b = first_state(subs[s])
        rts.append(r)
        offs.append(o)
        ths.append(t)
        lls.append(a)
        lps.append(b)
    n_tr = [len(r) for r in rts]
    out.update(pop_rt=np.concatenate(rts),
               pop_subject_off=np.concatenate([[0], np.cumsum(n_tr)]).astype(np.int64),
               pop_cell_off=np.stack(offs), pop_theta=np.stack(ths),
               pop_ll=np.stack(lls), pop_lp=np.stack(lps))
    phi_th, phi_ll, phi_lp = first_state(pop["phi"])
    out.update(phi_theta=phi_th, phi_ll=phi_ll, phi_lp=phi_lp,
               phi_pnames=np.array(pop["phi"].slot("pnames").value))
    priors = d["pop_priors"]
    for k, v in prior_spec(priors.slot("p_prior"), pnames).items():
        out["pop_prior_" + k] = v
    for k, v in prior_spec(priors.slot("h_prior"),
                           list(priors.slot("pnames").value)).items():
        out["hyper_prior_" + k] = v
    out["de_nparameter"] = np.int32(priors.slot("nparameter").value[0])
    return out


def main():
    for idx in (2, 3, 4, 5, 6):
        out = extract(idx)
        path = os.path.join(OUT_DIR, "lba_data%d.npz" % idx)
        np.savez_compressed(path, **out)
        print("%s: P=%d n_acc=%d sub trials=%d pop subjects=%d phi chains=%d -> %d KB"
              % (os.path.basename(path), len(out["design_pnames"]),
                 out["design_n_acc"], len(out["sub_rt"]), out["pop_theta"].shape[0],
                 out["phi_theta"].shape[1], os.path.getsize(path) // 1024))


if __name__ == "__main__":
    main()
