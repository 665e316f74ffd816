"""Benchmark of the DE-MCMC hot path on one MI355X.

Workload (default c3): hierarchical B x v LBA, 32 subjects x 768 trials,
13 subject parameters (26 population parameters), 39 chains, synthetic data
from the LBA generation rule. A step is one full DE-MCMC iteration (population
step + every subject's crossover / migration over all its trials).

Measured on the same inputs, all resident in HBM before the timed region:
  value          iterations/s of the batched path (mixed FP32/FP64 density
                 arithmetic with an error-model fallback to the FP64 formula,
                 FP64 sums and sampler state, hipGraph-replayed launches)
  plain_fp64     iterations/s of the plain FP64 GPU path (the reference's
                 arithmetic and sequential chain order, one launch per chain step)
  max_rel_dLL    max |LL_batched / LL_fp64 - 1| over every stored subject state
  roofline       of the subject sweep kernel (dominant): VALU issue bound,
                 algorithmic slots per evaluation x evaluations per launch /
                 launch time (HIP events on the launch stream)
  cpu_baseline   the FP64 oracle restatement of the reference sampler: one run
                 on one core (value), and as many independent runs as host
                 cores (at most 16) at once (aggregate, run-iterations/s)

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("DE-MCMC iter/s + trial-LL evals/s: batched FP32 vs plain FP64 GPU, "
          "|ΔLL|≤1e-6")

# Algorithmic VALU issue slots of one trial-likelihood evaluation (one trial
# x one proposal, all accumulators; a slot = one lane for one VALU cycle, a
# quarter-rate transcendental counts 4): per accumulator the shifted time,
# scale, two z-scores, two exponentials, two normal tails and the winner
# density or loser survivor, then the product, log and accumulate. About
# 86 arithmetic + 11 transcendental instructions with 2 accumulators,
# 170 + 21 with 4.
W_ALG = {2: 130.0, 4: 255.0}
# VALU lane-slots the subject sweep kernel actually issues per evaluation,
# everything included (precision work, indexing, staging, sums, proposals,
# the population workgroups, and the polling of items that wait for their
# subject's previous iteration): rocprofv3 SQ_INSTS_VALU x 64 / evaluations of
# a launch of 32 iterations (profiles/subject_sweep_kernel_<workload>.md).
# Keyed by (workload, chains per item, trials per item, launch mode, waves per
# SIMD of the kernel build), for the
# decompositions the autotuner picks on these workloads: a figure is reported
# only for the decomposition it was counted on.
W_ISSUED = {('c3', 1, 0, 4, 3): 735.6, ('c3', 1, 0, 3, 3): 889.1, ('c3', 2, 0, 4, 3): 607.4,
            ('c3', 2, 0, 3, 3): 683.7, ('c4', 4, 0, 3, 4): 519.7, ('c4', 4, 0, 3, 3): 504.5,
            ('c4', 4, 0, 4, 3): 473.8, ('c5', 17, 0, 3, 4): 664.3, ('c5', 17, 0, 3, 3): 661.6,
            ('c5', 17, 0, 4, 3): 659.9}
# HBM bytes per iteration of the subject sweep kernel at those decompositions:
# rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 KiB (factors of a 256 MiB copy
# kernel, same files) per launch, divided by the 32 iterations of a launch.
TRAFFIC = {('c3', 1, 0, 4, 3): 1.498e+06, ('c3', 1, 0, 3, 3): 1.238e+06,
           ('c3', 2, 0, 4, 3): 1.250e+06, ('c3', 2, 0, 3, 3): 1.231e+06,
           ('c4', 4, 0, 3, 4): 5.701e+06, ('c4', 4, 0, 3, 3): 5.456e+06,
           ('c4', 4, 0, 4, 3): 5.483e+06, ('c5', 17, 0, 3, 4): 2.212e+07,
           ('c5', 17, 0, 3, 3): 2.170e+07, ('c5', 17, 0, 4, 3): 2.186e+07}
# Iterations a bench session holds at least, whatever --steps asks: the
# library autotunes the work decomposition of sessions of 64 or more
# iterations, so the measured configuration never depends on the step count.
MIN_SESSION_ITERS = 128
# 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
VALU_PEAK_TSLOTS = 256 * 4 * 16 * 2.4e9 / 1e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--runs", type=int, default=1,
                    help="independent runs of the workload in one batch (Philox keys seed + r)")
    ap.add_argument("--plain-steps", type=int, default=0,
                    help="iterations of the plain FP64 path (0: 20, or 3 for c4 / c5)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-cores", type=int, default=0,
                    help="threads for the aggregate CPU baseline (0: the host cores, at most 16)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--profile-iters", type=int, default=32)
    ap.add_argument("--decomp", default="",
                    help="fixed work items 'k,t[,m[,w]]' (chains per item or -1, trials per item "
                         "or 0, launch mode 1 to 4 as make_config's launch_mode, waves per SIMD "
                         "of the subject kernel 3 or 4) instead of autotuning")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 256 MiB device-to-device copy (counter calibration)")
    return ap.parse_args()


def calibrate_copy():
    """A plain copy of known size: FETCH_SIZE / WRITE_SIZE of this kernel
    scale the counters of the sampler's kernels to bytes."""
    import torch
    n = 64 * 1024 * 1024   # 256 MiB of float32
    x = torch.ones(n, device="cuda", dtype=torch.float32)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}))


def host_cores():
    """Cores this process may run on, capped at 16 (the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, budget_s, cores):
    """Oracle (FP64, one thread per run) on the same workload, every job of it
    one after the other, for a bounded number of iterations. First one run on
    one core, then `cores` independent runs (seeds offset by the run index) on
    as many threads at once, the way the reference runs one sampler per forked
    process. Returns (per-core iter/s, n, seconds, aggregate iter/s, n_agg,
    wall seconds)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from ggdmc_amd import _lib
    om = O.Model(w.model.winner, w.model.par_src, w.model.par_const, w.model.npar)

    def opr(p):
        return O.Prior(p.dist, p.p0, p.p1, p.lower, p.upper)

    def one(j, n_iter, run):
        cfg = O.config(nmc=n_iter + 1, nchain=w.nchain, nparameter_de=j.nparameter_de,
                       sub_migration_prob=w.sub_migration_prob,
                       pop_migration_prob=w.pop_migration_prob, seed=j.seed + run)
        t = time.perf_counter()
        if j.kind == _lib.KIND_HIER:
            datas = j.subjects
            rt = np.concatenate([d.rt for d in datas])
            so = np.concatenate([[0], np.cumsum([len(d.rt) for d in datas])])
            co = np.stack([d.cell_off for d in datas])
            O.run_hier(cfg, om, opr(j.p_prior), opr(j.h_prior), rt, so, co, *j.phi_start,
                       np.stack([s[0] for s in j.subject_starts]),
                       np.stack([s[1] for s in j.subject_starts]),
                       np.stack([s[2] for s in j.subject_starts]))
        else:
            O.run_subject(cfg, om, opr(j.p_prior), j.subject.rt, j.subject.cell_off, *j.start)
        return time.perf_counter() - t

    def all_jobs(n_iter, run=0):
        return sum(one(j, n_iter, run) for j in w.jobs)

    t1 = all_jobs(1)
    n = max(2, int(budget_s / max(t1, 1e-6)))
    dt = all_jobs(n)
    # The oracle's C calls release the GIL, so threads run on separate cores.
    from concurrent.futures import ThreadPoolExecutor
    n_agg = max(2, int(n * 0.6))
    t = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:
        list(ex.map(lambda r: all_jobs(n_agg, r), range(cores)))
    wall = time.perf_counter() - t
    return n / dt, n, dt, cores * n_agg / wall, n_agg, wall


def main():
    a = parse()
    if a.calibrate_copy:
        calibrate_copy()
        return
    if a.gpus != 1:
        print("bench.py measures one GPU; the engine runs no multi-GPU configuration "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    import ggdmc_amd as g
    from ggdmc_amd import workloads
    if g.device_count() < 1:
        print("no gfx950 GPU visible", file=sys.stderr)
        sys.exit(1)

    t_build = time.time()
    w_one = workloads.make(a.workload)
    w = workloads.replicate(w_one, a.runs)
    n_acc = w.model.n_acc
    total = a.warmup + a.steps + a.profile_iters
    sess_iters = max(total, MIN_SESSION_ITERS)
    thin = max(1, sess_iters // 1000)
    nmc = sess_iters // thin + 2
    if a.decomp:
        dv = [int(v) for v in a.decomp.split(",")]
        k, t, m = dv[0], dv[1], (dv[2] if len(dv) > 2 else 1)
        wps = dv[3] if len(dv) > 3 else 3
        cfg = w.config(nmc=nmc, thin=thin, chains_per_item=k, trials_per_tile=t, launch_mode=m,
                       waves_per_simd=wps)
    else:
        cfg = w.config(nmc=nmc, thin=thin)
    sess = g.Session(cfg, w.model, w.jobs)
    kci, tpt, nitems, lmode = sess.decomposition()
    occ = sess.waves_per_simd()
    print("[bench] decomp %d,%d,%d (%d waves per SIMD)" % (kci, tpt, lmode, occ), file=sys.stderr, flush=True)
    print("[bench] workload %s: %s; %d subjects, %d trials; items: %s chains x %s trials, "
          "%d per launch, %s; setup %.1fs"
          % (w.name, w.description, w.n_subjects, w.n_trials,
             "all groups" if kci < 0 else kci, tpt or "all", nitems,
             {1: "a subject launch per chain group", 2: "one subject launch for every chain group",
              3: "one subject launch for every chain group and 32 iterations",
              4: "one subject launch for every chain group and 32 iterations, an item "
                 "holding its chain range of every group"}[lmode],
             time.time() - t_build),
          file=sys.stderr, flush=True)

    sess.advance(a.warmup)
    e0 = sess.trial_evals()
    t0 = time.perf_counter()
    ms = sess.advance(a.steps)
    wall = time.perf_counter() - t0
    e1 = sess.trial_evals()
    iters_per_s = a.steps / (ms / 1e3)
    evals_per_s = (e1 - e0) / (ms / 1e3)
    print("[bench] batched: %d iterations in %.1f ms (wall %.1f ms): %.0f iter/s, %.3g trial-evals/s"
          % (a.steps, ms, wall * 1e3, iters_per_s, evals_per_s), file=sys.stderr, flush=True)

    # dominant kernel, timed launch by launch with HIP events on its stream
    e2 = sess.trial_evals()
    k_ms, n_launch = sess.profile(a.profile_iters)
    e3 = sess.trial_evals()
    evals_per_launch = (e3 - e2) / max(n_launch, 1)
    iters_per_launch = a.profile_iters / max(n_launch, 1)
    achieved = evals_per_launch * W_ALG[n_acc] / (k_ms / 1e3) / 1e12
    key = (w.name, kci, tpt, lmode, occ)
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(VALU_PEAK_TSLOTS, 2),
                "unit": "Tslot/s", "frac": round(achieved / VALU_PEAK_TSLOTS, 4),
                "traffic": TRAFFIC[key] * iters_per_launch if key in TRAFFIC else None,
                "kernel": "subject_sweep_kernel", "iterations_per_launch": iters_per_launch,
                "kernel_ms": round(k_ms, 4), "evals_per_launch": evals_per_launch,
                "slots_per_eval": W_ALG[n_acc]}
    wi = W_ISSUED.get(key)
    if wi:
        issued = evals_per_launch * wi / (k_ms / 1e3) / 1e12
        roofline.update({"issued_slots_per_eval": wi, "issued": round(issued, 3),
                         "issued_frac": round(issued / VALU_PEAK_TSLOTS, 4)})
    if roofline["traffic"]:
        roofline["hbm_GBps"] = round(roofline["traffic"] / (k_ms / 1e3) / 1e9, 1)
    t_dl = time.perf_counter()
    res = sess.download()
    dl_ms = (time.perf_counter() - t_dl) * 1e3
    dl_bytes = sum(int(np.asarray(p.theta).nbytes + np.asarray(p.log_likelihoods).nbytes * 2)
                   for r in res for p in ([r["phi"]] if r["phi"] is not None else []) + r["subjects"])
    sess.close()
    print("[bench] samples download: %.1f MB in %.1f ms (%d stored columns)"
          % (dl_bytes / 1e6, dl_ms, nmc), file=sys.stderr, flush=True)

    # per-proposal |dLL|: the last stored state of every chain against the
    # FP64 path (every subject for up to 64 subjects, an even subsample
    # beyond); a non-finite ratio counts as a failure
    worst = 0.0
    n_cmp = 0
    pairs = []
    for job, r in zip(w.jobs, res):
        datas = [job.subject] if job.kind == g.KIND_SUBJECT else job.subjects
        pairs.extend(zip(datas, r["subjects"]))
    step = max(1, len(pairs) // 64)
    k_last = total // thin          # iterations 1..total ran; column i/thin holds iteration i
    for d, post in pairs[::step]:
        ref = g.loglike(w.model, d, post.theta[:, :, k_last].T, g.FP64_PLAIN)
        r = np.abs(post.log_likelihoods[:, k_last] / ref - 1)
        worst = max(worst, float(np.max(r)) if np.all(np.isfinite(r)) else float("inf"))
        n_cmp += len(ref)

    plain = None
    if not a.no_plain:
        n_plain = a.plain_steps or (3 if w.name in ("c4", "c5") else 20)
        pcfg = w.config(nmc=n_plain + 3, precision=g.FP64_PLAIN)
        ps = g.Session(pcfg, w.model, w.jobs)
        ps.advance(2)
        pms = ps.advance(n_plain)
        ps.close()
        plain = n_plain / (pms / 1e3)
        print("[bench] plain FP64: %d iterations in %.1f ms: %.1f iter/s"
              % (n_plain, pms, plain), file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu:
        cores = a.cpu_cores or host_cores()
        cps, n_cpu, dt, agg, n_agg, wall = cpu_baseline(w_one, a.cpu_seconds, cores)
        cpu = {"value": round(cps, 4), "unit": "iter/s", "cores": 1, "kind": "port",
               "sample": "%d iterations of the same %s workload (all %d runs) in %.1f s "
                         "(FP64 oracle, one thread)" % (n_cpu, w_one.name, len(w_one.jobs), dt),
               "cpu_model": cpu_model(),
               "aggregate": {"value": round(agg, 4), "unit": "run-iter/s", "cores": cores,
                             "sample": "%d independent runs x %d iterations each, one per "
                                       "thread, in %.1f s wall" % (cores, n_agg, wall)}}
        print("[bench] cpu oracle: %s" % cpu["sample"], file=sys.stderr, flush=True)
        print("[bench] cpu oracle, %d cores: %s: %.3f run-iter/s (%.3f per core)"
              % (cores, cpu["aggregate"]["sample"], agg, agg / cores),
              file=sys.stderr, flush=True)

    out = {
        "metric": METRIC, "value": round(iters_per_s, 2), "unit": "iter/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms / a.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32/fp64 mixed (fp64 exp, sums, state)", "data": "synthetic",
        "config": {"workload": "%s: %s" % (w.name, w.description), "subjects": w.n_subjects,
                   "trials": w.n_trials, "chains": w.nchain, "sweep_groups": 2,
                   "chains_per_item": kci, "trials_per_item": tpt, "items_per_launch": nitems,
                   "launch_mode": lmode, "waves_per_simd": occ, "runs": a.runs},
        "run_iter_per_s": round(iters_per_s * a.runs, 2),
        "trial_evals_per_s": round(evals_per_s, 1),
        "plain_fp64_iter_per_s": round(plain, 3) if plain else None,
        "speedup_vs_plain_fp64": round(iters_per_s / plain, 2) if plain else None,
        "speedup_vs_cpu": round(iters_per_s / cpu["value"], 1) if cpu else None,
        "speedup_vs_cpu_all_cores": (round(iters_per_s * a.runs / cpu["aggregate"]["value"], 1)
                                     if cpu else None),
        "max_rel_dLL": worst, "dLL_states_checked": n_cmp,
        "download": {"ms": round(dl_ms, 2), "MB": round(dl_bytes / 1e6, 2), "columns": nmc},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
