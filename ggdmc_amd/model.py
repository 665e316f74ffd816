"""Factorial LBA design -> flattened design of the C ABI.

Counterpart of ggdmcModel::BuildModel / BuildDMI as used by the reference's
data generators (tests/testthat/Group1/data/6_lba_Bv_model.r:28-39,94): a
design is given by
  p_map        family -> factors it varies with ("1" = none; "M" = the match
               factor, true when the accumulator is the stimulus' match)
  match_map    {"M": {stimulus level: matching accumulator}}
  factors      {factor: levels}
  constants    {parameter-condition name: value}
  accumulators response names, in model order.
Cells are every combination of factor levels and response, named
"<levels>.<response>" and sorted; free parameter names are the sorted
parameter-condition names that are not constants. The winner of a cell is its
response; dmi@node_1_index lists it first.
"""
import itertools

import numpy as np

from .runs import LBAModel, SubjectData

FAMILIES = ("A", "B", "mean_v", "sd_v", "st0", "t0")


def _stimulus_factor(match_map, factors):
    keys = set(next(iter(match_map.values())).keys())
    for f, levels in factors.items():
        if set(levels) == keys:
            return f
    raise ValueError("no factor has the levels of match_map")


def build_model(p_map, match_map, factors, constants, accumulators):
    """Returns an LBAModel (flattened design) plus its cell names and pnames."""
    accumulators = list(accumulators)
    fnames = list(factors)
    stim = _stimulus_factor(match_map, factors) if match_map else None
    mm = next(iter(match_map.values())) if match_map else {}
    for fam in FAMILIES:
        if fam not in p_map:
            raise ValueError("p_map lacks family " + fam)

    def pcond_name(fam, levels, acc):
        spec = p_map[fam]
        if list(spec) == ["1"]:
            return fam
        tags = []
        for f in spec:
            if f == "M":
                tags.append("true" if mm[levels[stim]] == acc else "false")
            else:
                tags.append(levels[f])
        return fam + "." + ".".join(tags)

    combos = list(itertools.product(*[factors[f] for f in fnames]))
    cells = sorted(".".join(list(c) + [r]) for c in combos for r in accumulators)
    all_pc = set()
    table = {}
    for name in cells:
        parts = name.split(".")
        levels = dict(zip(fnames, parts[:-1]))
        for j, acc in enumerate(accumulators):
            for fam in FAMILIES:
                pc = pcond_name(fam, levels, acc)
                table[(name, j, fam)] = pc
                all_pc.add(pc)
    pnames = sorted(pc for pc in all_pc if pc not in constants)
    n_cell, n_acc = len(cells), len(accumulators)
    par_src = np.full((n_cell, n_acc, 6), -1, np.int32)
    par_const = np.zeros((n_cell, n_acc, 6))
    winner = np.zeros(n_cell, np.int32)
    for c, name in enumerate(cells):
        winner[c] = accumulators.index(name.split(".")[-1])
        for j in range(n_acc):
            for f, fam in enumerate(FAMILIES):
                pc = table[(name, j, fam)]
                if pc in constants:
                    par_const[c, j, f] = constants[pc]
                else:
                    par_src[c, j, f] = pnames.index(pc)
    return LBAModel(winner, par_src, par_const, pnames, cells)


def build_dmi(model, trials):
    """BuildDMI: trials is a list of (cell name or (levels..., response), rt);
    returns the subject's cell-sorted SubjectData."""
    index = {n: c for c, n in enumerate(model.cell_names)}
    cells = [[] for _ in range(model.n_cell)]
    for key, rt in trials:
        name = key if isinstance(key, str) else ".".join(key)
        cells[index[name]].append(float(rt))
    return SubjectData.from_cells([np.array(c) for c in cells])
